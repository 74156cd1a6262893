"""ctypes bindings of the CPU oracle (oracle/liborc.so) and of the reference's
own CPU engine built from its sources (oracle/_ref/libref.so). Test
infrastructure only."""
from __future__ import annotations

import contextlib
import ctypes
import os
from pathlib import Path

import numpy as np

from raymotion_amd._capi import GBUFFER_DTYPE, RESERVOIR_DTYPE, Frame, SceneDesc, ptr

ROOT = Path(__file__).resolve().parent.parent
ORC_PATH = Path(os.environ.get("RAYMOTION_ORACLE_LIB") or ROOT / "oracle" / "liborc.so")
REF_PATH = ROOT / "oracle" / "_ref" / "libref.so"
THREADS = min(16, os.cpu_count() or 1)

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t


class OrcDebug(ctypes.Structure):
    _fields_ = [("gbuf", _P), ("rinit", _P), ("rmerged", _P), ("shade", _P), ("path", _P), ("lanes", _P),
                ("has_history", _P), ("stats", _P)]


_orc = None
_ref = None


def orc() -> ctypes.CDLL:
    global _orc
    if _orc is None:
        if not ORC_PATH.exists():
            raise RuntimeError(f"{ORC_PATH} missing: run __graft_entry__.build()")
        lib = ctypes.CDLL(str(ORC_PATH))
        sigs = [
            ("orc_history_create", _P, []),
            ("orc_history_destroy", None, [_P]),
            ("orc_render_frame", ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Frame), ctypes.c_uint64, _P,
                                                ctypes.c_int, _P, _P, ctypes.POINTER(OrcDebug)]),
            ("orc_render_reference_f64", ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Frame),
                                                        ctypes.c_uint64, ctypes.c_int, _P]),
            ("orc_trace", ctypes.c_int, [ctypes.POINTER(SceneDesc), _P, _P, _P, _SZ, ctypes.c_int, ctypes.c_int, _P, _P,
                                         _P, _P, _P]),
            ("orc_trace_brute", ctypes.c_int, [ctypes.POINTER(SceneDesc), _P, _P, _SZ, ctypes.c_int, _P, _P]),
            ("orc_rng_draws", None, [ctypes.c_uint64, ctypes.c_int, _P, _P]),
            ("orc_render_adjoint", ctypes.c_longlong, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Frame),
                                                       ctypes.c_uint64, _P, _P, _P, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_int, _P]),
            ("orc_bsdf_grad_f64", None, [_P, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P, _P, _P]),
            ("orc_sincos", None, [_P, _SZ, _P, _P]),
            ("orc_check_div_pi", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
            ("orc_check_exp", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _P]),
            ("orc_check_atan", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _P]),
            ("orc_set_wide_tree", None, [_P, _SZ, _P]),
            ("orc_trace_wide", ctypes.c_int, [_P, _SZ, _P, _P, _P, _P, _SZ, ctypes.c_int, ctypes.c_int, _P, _P,
                                              _P]),
            ("orc_restir_merge", ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Frame), ctypes.c_uint64,
                                                _P, _P, _P, _P, _P, ctypes.c_int, _P]),
            ("orc_denoise_image", ctypes.c_int, [ctypes.POINTER(Frame), _P, _P, _P, _P, _P, ctypes.c_int, _P]),
        ]
        for name, res, args in sigs:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _orc = lib
    return _orc


def ref() -> ctypes.CDLL | None:
    """The reference engine library, or None where it was not built."""
    global _ref
    if _ref is None and REF_PATH.exists():
        lib = ctypes.CDLL(str(REF_PATH))
        sigs = [
            ("ref_scene_create", _P, []), ("ref_scene_destroy", None, [_P]), ("ref_scene_clear", None, [_P]),
            ("ref_add_triangles", None, [_P, _SZ, _P, _P, _P, _P, _P]),
            ("ref_build_bvh", None, [_P, ctypes.c_int]), ("ref_refit_bvh", None, [_P]),
            ("ref_counts", None, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
            ("ref_nodes", None, [_P, _P, _P, _P]),
            ("ref_precompute", None, [_P, _P, _P, _P, _P, _P]),
            ("ref_camera", None, [_P, _P, ctypes.c_double, _P]),
        ]
        for name, res, args in sigs:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _ref = lib
    return _ref


class History:
    def __init__(self):
        self.h = orc().orc_history_create()

    def __del__(self):
        if getattr(self, "h", None):
            orc().orc_history_destroy(self.h)
            self.h = None


def render_frame(exported, frame: Frame, seed: int, history: History | None = None, debug: bool = False,
                 threads: int = THREADS):
    """One oracle frame: returns (rgb8, hdr, info) where info holds the
    intermediates when debug is set."""
    d = exported.desc()
    px = frame.width * frame.height
    rgb = np.zeros(px * 3, np.uint8)
    hdr = np.zeros(px * 3, np.float32)
    info = {}
    dbg = OrcDebug()
    path = ctypes.c_int32(-1)
    lanes = ctypes.c_int32(0)
    hh = ctypes.c_int32(0)
    stats = np.zeros(7, np.uint64)
    dbg.path = ctypes.addressof(path)
    dbg.lanes = ctypes.addressof(lanes)
    dbg.has_history = ctypes.addressof(hh)
    dbg.stats = ptr(stats)
    if debug:
        info["gbuf"] = np.zeros(px, GBUFFER_DTYPE)
        info["rinit"] = np.zeros(px, RESERVOIR_DTYPE)
        info["rmerged"] = np.zeros(px, RESERVOIR_DTYPE)
        info["shade"] = np.zeros(px * 3, np.float32)
        dbg.gbuf = ptr(info["gbuf"])
        dbg.rinit = ptr(info["rinit"])
        dbg.rmerged = ptr(info["rmerged"])
        dbg.shade = ptr(info["shade"])
    rc = orc().orc_render_frame(ctypes.byref(d), ctypes.byref(frame), seed, history.h if history else None, threads,
                                ptr(rgb), ptr(hdr), ctypes.byref(dbg))
    if rc:
        raise RuntimeError(f"orc_render_frame failed ({rc})")
    info.update(path=path.value, lanes=lanes.value, has_history=hh.value, rays=int(stats[0]),
                node_visits=int(stats[1]), tri_tests=int(stats[2]),
                pass_rays={"megakernel": int(stats[3]), "generate": int(stats[4]), "spatial": int(stats[5]),
                           "shade": int(stats[6])})
    return rgb, hdr, info


def restir_merge(exported, frame: Frame, seed: int, gbuf, rinit, hist_gbuf=None, hist_res=None, prev_camera=None,
                 threads: int = THREADS) -> np.ndarray:
    """The oracle's restir_spatial alone over given G-buffer and reservoirs."""
    d = exported.desc()
    g = np.ascontiguousarray(gbuf, GBUFFER_DTYPE)
    ri = np.ascontiguousarray(rinit, RESERVOIR_DTYPE)
    hg = None if hist_gbuf is None else np.ascontiguousarray(hist_gbuf, GBUFFER_DTYPE)
    hr = None if hist_res is None else np.ascontiguousarray(hist_res, RESERVOIR_DTYPE)
    out = np.zeros(frame.width * frame.height, RESERVOIR_DTYPE)
    cam = None if prev_camera is None else ctypes.addressof(prev_camera)
    rc = orc().orc_restir_merge(ctypes.byref(d), ctypes.byref(frame), seed, ptr(g), ptr(ri), ptr(hg), ptr(hr), cam,
                                threads, ptr(out))
    if rc:
        raise RuntimeError("orc_restir_merge failed")
    return out


def denoise_image(frame: Frame, color, gbuf, hist_color=None, hist_gbuf=None, prev_camera=None,
                  threads: int = THREADS) -> np.ndarray:
    """The oracle's temporal + a-trous passes alone (W*H*3 out)."""
    col = np.ascontiguousarray(color, np.float32).reshape(-1)
    g = np.ascontiguousarray(gbuf, GBUFFER_DTYPE)
    hc = None if hist_color is None else np.ascontiguousarray(hist_color, np.float32).reshape(-1)
    hg = None if hist_gbuf is None else np.ascontiguousarray(hist_gbuf, GBUFFER_DTYPE)
    out = np.zeros(col.size, np.float32)
    cam = None if prev_camera is None else ctypes.addressof(prev_camera)
    rc = orc().orc_denoise_image(ctypes.byref(frame), ptr(col), ptr(g), ptr(hc), ptr(hg), cam, threads, ptr(out))
    if rc:
        raise RuntimeError("orc_denoise_image failed")
    return out


def render_reference_f64(exported, frame: Frame, seed: int, threads: int = THREADS) -> np.ndarray:
    d = exported.desc()
    out = np.zeros(frame.width * frame.height * 3, np.float64)
    orc().orc_render_reference_f64(ctypes.byref(d), ctypes.byref(frame), seed, threads, ptr(out))
    return out


def render_adjoint(exported, frame: Frame, seed: int, d_hdr, tri_group, group_emission, precision: int = 64,
                   threads: int = THREADS) -> np.ndarray:
    """The oracle's path-replay adjoint: (ngroup, 5) gradients of sum(d_hdr * hdr)."""
    d = exported.desc()
    dh = np.ascontiguousarray(d_hdr, np.float32).reshape(-1)
    tg = np.ascontiguousarray(tri_group, np.int32)
    ge = np.ascontiguousarray(group_emission, np.float32).reshape(-1, 3)
    grad = np.zeros((len(ge), 5), np.float64)
    bad = orc().orc_render_adjoint(ctypes.byref(d), ctypes.byref(frame), seed, ptr(dh), ptr(tg), ptr(ge), len(ge),
                                   precision, threads, ptr(grad))
    if bad < 0:
        raise RuntimeError("orc_render_adjoint: frame not covered by the adjoint")
    if bad > 0:
        raise RuntimeError(f"orc_render_adjoint: {bad} samples replayed differently")
    return grad


def bsdf_grad_f64(wi, wo, n, alpha2, metallic, albedo):
    """FP64 opaque BSDF value, d value / d albedo (own channel) and d value / d alpha2."""
    arr = [np.ascontiguousarray(v, np.float64) for v in (wi, wo, n, albedo)]
    val, da, dl = np.zeros(3), np.zeros(3), np.zeros(3)
    orc().orc_bsdf_grad_f64(ptr(arr[0]), ptr(arr[1]), ptr(arr[2]), float(alpha2), float(metallic), ptr(arr[3]),
                            ptr(val), ptr(da), ptr(dl))
    return val, da, dl


def trace(exported, origins, dirs, tmax=None, any_hit=False, threads: int = THREADS):
    d = exported.desc()
    o = np.ascontiguousarray(origins, np.float32)
    dd = np.ascontiguousarray(dirs, np.float32)
    n = len(o)
    tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
    tri = np.empty(n, np.int32)
    t = np.empty(n, np.float32); u = np.empty(n, np.float32); v = np.empty(n, np.float32)
    stats = np.zeros(2, np.uint64)
    orc().orc_trace(ctypes.byref(d), ptr(o), ptr(dd), ptr(tm), n, int(any_hit), threads, ptr(tri), ptr(t), ptr(u),
                    ptr(v), ptr(stats))
    return tri, t, u, v, stats


@contextlib.contextmanager
def wide_tree(nodes, tris):
    """Inside the block, FP32 oracle frames (render_frame, restir_merge,
    render_adjoint) traverse the backend's own 4-wide tree (Backend.read_tree)
    with the backend's traversal instead of the scene's binary tree: the same
    boxes and slab arithmetic as the GPU, so also the same outcome for rays
    that graze a box edge, where tight boxes make the result depend on the
    tree."""
    nd = np.ascontiguousarray(nodes, np.float32)
    tr = np.ascontiguousarray(tris, np.float32)
    orc().orc_set_wide_tree(ptr(nd), len(nd), ptr(tr))
    try:
        yield
    finally:
        orc().orc_set_wide_tree(None, 0, None)


def trace_wide(nodes, tris, origins, dirs, tmax=None, any_hit=False, threads: int = THREADS):
    """Host walk of the backend's own 4-wide tree (Backend.read_tree): hit ids,
    distances and the [node visits, triangle tests] it makes."""
    nd = np.ascontiguousarray(nodes, np.float32)
    tr = np.ascontiguousarray(tris, np.float32)
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = len(o)
    tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
    tri = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    stats = np.zeros(2, np.uint64)
    orc().orc_trace_wide(ptr(nd), len(nd), ptr(tr), ptr(o), ptr(d), ptr(tm), n, int(any_hit), threads, ptr(tri),
                         ptr(t), ptr(stats))
    return tri, t, stats


def trace_brute(exported, origins, dirs, threads: int = THREADS):
    d = exported.desc()
    o = np.ascontiguousarray(origins, np.float32)
    dd = np.ascontiguousarray(dirs, np.float32)
    n = len(o)
    tri = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    orc().orc_trace_brute(ctypes.byref(d), ptr(o), ptr(dd), n, threads, ptr(tri), ptr(t))
    return tri, t


def rng_draws(seed: int, n: int):
    out = np.empty(n, np.float32)
    states = np.empty(n + 1, np.uint32)
    orc().orc_rng_draws(seed, n, ptr(out), ptr(states))
    return out, states


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    orc().orc_sincos(ptr(x), len(x), ptr(s), ptr(c))
    return s, c


def div_pi_mismatches(lo_bits: int, hi_bits: int, threads: int = 8) -> int:
    """Floats with |x|'s bits in [lo_bits, hi_bits), both signs, for which the
    shared division by pi differs from the IEEE quotient x / pi."""
    return int(orc().orc_check_div_pi(lo_bits, hi_bits, threads))


def exp_ulp(lo_bits: int, hi_bits: int, threads: int = 8):
    """Floats with |x|'s bits in [lo_bits, hi_bits), both signs: the largest
    difference in ulp between the shared exponential and libm's double exp
    rounded to float (2^31 for a NaN on one side only), and how many differ."""
    n = ctypes.c_uint64(0)
    worst = int(orc().orc_check_exp(lo_bits, hi_bits, threads, ctypes.addressof(n)))
    return worst, int(n.value)


def atan_ulp(lo_bits: int, hi_bits: int, threads: int = 8):
    """Floats x >= 0 with bits in [lo_bits, hi_bits): the largest difference
    in ulp between the shared arctangent and libm's double atan rounded to
    float (2^31 for a NaN on one side only, or atan(-x) != -atan(x)), and how
    many differ."""
    n = ctypes.c_uint64(0)
    worst = int(orc().orc_check_atan(lo_bits, hi_bits, threads, ctypes.addressof(n)))
    return worst, int(n.value)


class RefScene:
    """The reference engine's Scene (oracle/_ref/libref.so)."""

    def __init__(self):
        self.lib = ref()
        self.h = self.lib.ref_scene_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_scene_destroy(self.h)
            self.h = None

    def add_triangles(self, pos, mat, normals=None, uv=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 9)
        n = len(pos)
        mat = np.ascontiguousarray(np.broadcast_to(mat, (n, 11)), np.float64)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(n, 9)
        uvs = None if uv is None else np.ascontiguousarray(uv, np.float64).reshape(n, 6)
        idx = np.empty(n, np.int32)
        self.lib.ref_add_triangles(self.h, n, ptr(pos), ptr(nrm), ptr(uvs), ptr(mat), ptr(idx))
        return idx

    def clear(self):
        self.lib.ref_scene_clear(self.h)

    def build_bvh(self, leaf=8):
        self.lib.ref_build_bvh(self.h, leaf)

    def refit_bvh(self):
        self.lib.ref_refit_bvh(self.h)

    def counts(self):
        a, b, c = _SZ(), _SZ(), _SZ()
        self.lib.ref_counts(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def nodes(self):
        ntri, nnode, _ = self.counts()
        boxes = np.empty((nnode, 6)); links = np.empty((nnode, 4), np.int32)
        bvh_tri = np.empty(ntri if nnode else 0, np.int32)
        if nnode:
            self.lib.ref_nodes(self.h, ptr(boxes), ptr(links), ptr(bvh_tri))
        return boxes, links, bvh_tri

    def precompute(self):
        ntri, _, nlight = self.counts()
        n = np.empty((ntri, 3)); nn = np.empty((ntri, 3)); denom = np.empty(ntri); cdf = np.empty(nlight)
        total = ctypes.c_double()
        self.lib.ref_precompute(self.h, ptr(n), ptr(nn), ptr(denom), ptr(cdf), ctypes.byref(total))
        return n, nn, denom, cdf, total.value


def ref_camera(pos, rot, fov):
    out = np.empty(13)
    p = np.ascontiguousarray(pos, np.float64)
    r = np.ascontiguousarray(rot, np.float64)
    ref().ref_camera(ptr(p), ptr(r), float(fov), ptr(out))
    return out
