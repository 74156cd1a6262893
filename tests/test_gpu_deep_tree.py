"""Trees whose traversal stack bound exceeds the 64 LDS entries per lane run
on the deep kernel instance (the entries past 64 in per-lane scratch memory),
and very deep caller trees collapse without recursion. The reference's
stackless preorder walk (pathtrace.metal:119-163) renders any tree; these
trees come from tests/deep_trees.py."""
import numpy as np
import pytest

import deep_trees
import oracle_bindings as ob
from helpers import frame, rel_err
from test_gpu_parity import check_pass_rays

pytestmark = pytest.mark.gpu


def _trace_checks(backend, e, depth, n=1 << 14):
    o, d = deep_trees.rays(depth, n, 11)
    g_tri, g_t, _, _ = backend.trace(o, d)
    st = backend.stats()
    nodes, tris, need = backend.read_tree()
    b_tri, b_t = ob.trace_brute(e, o, d)
    w_tri, w_t, walk = ob.trace_wide(nodes, tris, o, d)
    print(f"depth {depth}: {len(nodes)} wide nodes, stack bound {need}, GPU {st.node_visits[4]} node visits "
          f"(host walk {int(walk[0])}), hit fraction {np.mean(g_tri >= 0):.3f}", flush=True)
    assert np.array_equal(g_tri, b_tri) and np.array_equal(g_t[g_tri >= 0], b_t[b_tri >= 0])
    assert np.array_equal(g_tri, w_tri)
    assert st.node_visits[4] == int(walk[0]) and st.tri_tests[4] == int(walk[1])
    return need


def test_comb_beyond_lds_stack(backend):
    """Stack bound 120 (> 64): closest hits equal brute force, node visits
    equal the host walk of the same wide tree; rays along -x push three
    entries per level, so they reach the scratch part of the stack."""
    e = deep_trees.comb(120)
    backend.set_scene(e)
    need = _trace_checks(backend, e, 120)
    assert 64 < need <= 256


def test_comb_beyond_lds_stack_frames(backend):
    """Megakernel and ReSTIR frames of the deep comb equal the oracle."""
    depth = 120
    e = deep_trees.comb(depth)
    backend.set_scene(e)
    assert backend.read_tree()[2] > 64
    cam = deep_trees.camera(depth, 48)
    for f in (frame(48, 40, 4, 5, cam=cam), frame(48, 40, 4, 4, cam=cam, candidates=8, spatial=2, radius=8.0)):
        g_rgb, g_hdr = backend.render(f, 99)
        c_rgb, c_hdr, info = ob.render_frame(e, f, 99)
        err = rel_err(g_hdr, c_hdr)
        print(f"path {info['path']}: max rel {err.max():.3g}, exact {np.mean(g_hdr == c_hdr):.4f}, "
              f"mean {c_hdr.mean():.4f}", flush=True)
        assert c_hdr.max() > 0
        assert err.max() <= 1e-3
        check_pass_rays(backend.stats(), info)


def test_too_deep_tree_rejected(backend):
    """A stack bound beyond the deep instance's 256 entries is an error, not
    a fault; the GPU tree build of the same triangles renders."""
    import raymotion_amd as rm
    e = deep_trees.comb(400)
    with pytest.raises(rm.BackendError, match="BVH too deep"):
        backend.set_scene(e)
    backend.set_scene(e, gpu_tree=True, rebuild=True)
    o, d = deep_trees.rays(400, 4096, 5)
    g_tri, _, _, _ = backend.trace(o, d)
    b_tri, _ = ob.trace_brute(e, o, d)
    assert np.array_equal(g_tri, b_tri)


def test_caterpillar_very_deep(backend):
    """A 60000-level caller tree (20000 wide levels) packs and traverses."""
    depth = 60000
    e = deep_trees.caterpillar(depth)
    backend.set_scene(e)
    need = _trace_checks(backend, e, depth, n=4096)
    assert need <= 1
