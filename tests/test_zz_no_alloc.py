"""rm_trace_device, the entry point that allocates nothing, against rm_trace
on the same rays: two calls in a row (the second smaller), bit-identical hit
ids and t/u/v, through the C-ABI with caller-allocated device buffers."""
import ctypes

import numpy as np
import pytest

import raymotion_amd as rm
from helpers import cornell
from raymotion_amd import rng as np_rng
from raymotion_amd._capi import lib

pytestmark = pytest.mark.gpu


def _device_trace(backend, o, d, tmax):
    import torch
    dev = torch.device("cuda:0")
    n = len(o)
    to = torch.from_numpy(np.ascontiguousarray(o, np.float32)).to(dev)
    td = torch.from_numpy(np.ascontiguousarray(d, np.float32)).to(dev)
    tt = torch.from_numpy(np.ascontiguousarray(tmax, np.float32)).to(dev)
    tri = torch.full((n,), -7, dtype=torch.int32, device=dev)
    tuv = torch.full((n, 4), -7.0, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    rc = lib().rm_trace_device(backend._h, to.data_ptr(), td.data_ptr(), tt.data_ptr(), n, 0, tri.data_ptr(),
                               tuv.data_ptr())
    assert rc == 0, lib().rm_last_error(backend._h).decode()
    torch.cuda.synchronize()
    return tri.cpu().numpy(), tuv.cpu().numpy()


def test_trace_device_matches_trace():
    b = rm.Backend(0)
    try:
        b.set_scene(cornell().export())
        for n in (1 << 18, 1000):  # second call smaller: the scratch keeps the first call's counters
            o, d = np_rng.random_rays(n, seed=n)
            tmax = np.full(n, 0.9, np.float32)
            tri_ref, t_ref, u_ref, v_ref = b.trace(o, d, tmax)
            st_ref = b.stats()
            tri, tuv = _device_trace(b, o, d, tmax)
            st = b.stats()
            print(f"rm_trace_device n={n}: hits {int((tri >= 0).sum())}", flush=True)
            assert np.array_equal(tri, tri_ref)
            hit = tri >= 0
            for k, ref in ((0, t_ref), (1, u_ref), (2, v_ref)):
                assert np.array_equal(tuv[hit, k], ref[hit])
            assert st.rays[4] == st_ref.rays[4] == n
            assert st.node_visits[4] == st_ref.node_visits[4] and st.tri_tests[4] == st_ref.tri_tests[4]
    finally:
        b.close()


def test_trace_device_rejects_host_memory():
    b = rm.Backend(0)
    try:
        b.set_scene(cornell().export())
        o = np.zeros((4, 3), np.float32)
        out = np.zeros(4, np.int32)
        tuv = np.zeros((4, 4), np.float32)
        rc = lib().rm_trace_device(b._h, o.ctypes.data, o.ctypes.data, None, 4, 0, out.ctypes.data, tuv.ctypes.data)
        assert rc == rm._capi.RM_ERR_INVALID
        assert b"device memory" in lib().rm_last_error(b._h)
    finally:
        b.close()
