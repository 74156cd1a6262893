"""GPU parity: the HIP kernels against the CPU oracle on identical inputs.

Bar: bit-identical closest-hit triangle ids (distances within 1e-6 relative);
per-pixel FP32 radiance within 1e-3 relative (floor 1e-3 absolute) and
identical 8-bit output wherever the radiance agrees; with the oracle on the
backend's own 4-wide tree (adaptive lanes, hue shader, ReSTIR pipeline),
radiance and 8-bit output bit-identical. Sizes are ones the oracle finishes
in seconds.
"""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import cornell, cornell_frame, frame, metal_test_scene, rel_err
from kat_cases import lit_textured_scene
from raymotion_amd import scenes
from raymotion_amd.rng import random_rays

pytestmark = pytest.mark.gpu


def _trace_parity(backend, exported, o, d, tmax=None, any_hit=False):
    backend.set_scene(exported)
    g_tri, g_t, _, _ = backend.trace(o, d, tmax, any_hit)
    c_tri, c_t, _, _, _ = ob.trace(exported, o, d, tmax, any_hit)
    if any_hit:
        return (g_tri >= 0) != (c_tri >= 0), None
    bad = g_tri != c_tri
    hit = (c_tri >= 0) & ~bad
    t_rel = np.abs(g_t[hit].astype(np.float64) - c_t[hit]) / np.maximum(np.abs(c_t[hit]), 1e-30)
    return bad, t_rel


def test_trace_cornell_2p20_bitexact(backend):
    e = cornell().export()
    o, d = random_rays(1 << 20, 0)
    bad, t_rel = _trace_parity(backend, e, o, d)
    print(f"cornell 2^20 rays: id mismatches {bad.sum()}, max t rel {t_rel.max():.3g}")
    assert bad.sum() == 0
    assert t_rel.max() <= 1e-6


@pytest.mark.parametrize("scene", ["cornell", "mixed"])
def test_trace_counters_match_host_walk(backend, scene):
    """Node visits and triangle tests the GPU counts (the roofline's traversal
    bytes) equal a host walk of the same resident 4-wide tree over the same
    rays, and so do the hits."""
    if scene == "cornell":
        e = cornell().export()
        o, d = random_rays(1 << 20, 0)
        backend.set_scene(e)
    else:
        s = scenes.mixed_spheres(seed=0, n_spheres=16)
        e = s.export()
        o, d = random_rays(1 << 17, 3, lo=(-5, -5, -5), hi=(5, 5, 5))
        backend.set_scene(e, gpu_tree=True, rebuild=True)
    g_tri, g_t, _, _ = backend.trace(o, d)
    st = backend.stats()
    nodes, tris, _ = backend.read_tree()
    c_tri, c_t, walk = ob.trace_wide(nodes, tris, o, d)
    print(f"{scene}: GPU {st.node_visits[4]} node visits, {st.tri_tests[4]} triangle tests; "
          f"host walk {int(walk[0])}, {int(walk[1])}", flush=True)
    assert np.array_equal(g_tri, c_tri) and np.array_equal(g_t[g_tri >= 0], c_t[c_tri >= 0])
    assert st.rays[4] == len(o)
    assert st.node_visits[4] == int(walk[0]) and st.tri_tests[4] == int(walk[1])


def test_trace_cornell_anyhit(backend):
    e = cornell().export()
    o, d = random_rays(1 << 18, 7)
    tmax = np.random.default_rng(1).uniform(0.01, 1.5, len(o)).astype(np.float32)
    bad, _ = _trace_parity(backend, e, o, d, tmax=tmax, any_hit=True)
    assert bad.sum() == 0


def test_trace_mixed_scene_bitexact(backend):
    s = scenes.mixed_spheres(seed=0, n_spheres=16)
    s.build_bvh()
    e = s.export()
    o, d = random_rays(1 << 17, 3, lo=(-5, -5, -5), hi=(5, 5, 5))
    bad, t_rel = _trace_parity(backend, e, o, d)
    print(f"mixed 2^17 rays: id mismatches {bad.sum()}, max t rel {t_rel.max():.3g}")
    assert bad.sum() == 0
    assert t_rel.max() <= 1e-6


def _compare_render(backend, exported, f, seed, tol=1e-3):
    backend.set_scene(exported)
    g_rgb, g_hdr = backend.render(f, seed)
    c_rgb, c_hdr, info = ob.render_frame(exported, f, seed)
    err = rel_err(g_hdr, c_hdr)
    exact = np.mean(g_hdr == c_hdr)
    print(f"path {info['path']} lanes {info['lanes']}: max rel {err.max():.3g}, exact fraction {exact:.6f}, "
          f"rgb mismatches {(g_rgb != c_rgb).sum()}")
    assert err.max() <= tol
    # 8-bit output must agree wherever the radiance agrees bit for bit.
    same = (g_hdr == c_hdr)
    assert np.all(g_rgb[same] == c_rgb[same])
    check_pass_rays(backend.stats(), info)
    return info


def check_pass_rays(st, info):
    """The GPU's per-pass ray counters (which the throughput figures divide by)
    equal the rays the oracle traced in the same passes."""
    gpu = {"megakernel": st.rays[0], "generate": st.rays[1], "spatial": st.rays[2], "shade": st.rays[3]}
    assert gpu == info["pass_rays"], (gpu, info["pass_rays"])
    assert st.rays[5] == info["rays"]


def test_render_cornell_c0(backend):
    e = cornell().export()
    info = _compare_render(backend, e, cornell_frame(256, 256, 4, 5), 12345)
    assert info["path"] == 0 and info["lanes"] == 1


@pytest.mark.parametrize("spp,lanes", [(8, 8), (32, 32)])
def test_render_cornell_lanes(backend, spp, lanes):
    e = cornell().export()
    info = _compare_render(backend, e, cornell_frame(48, 40, spp, 5), 777)
    assert info["lanes"] == lanes


def _exact_on_backend_tree(backend, exported, f, seed):
    """Radiance and 8-bit output bit-identical with the oracle traversing the
    backend's own 4-wide tree (the scene is already set)."""
    g_rgb, g_hdr = backend.render(f, seed)
    nodes, tris, _ = backend.read_tree()
    with ob.wide_tree(nodes, tris):
        c_rgb, c_hdr, _ = ob.render_frame(exported, f, seed)
    assert np.array_equal(np.asarray(g_hdr, np.float32).view(np.uint32).reshape(-1),
                          np.ascontiguousarray(c_hdr, np.float32).view(np.uint32).reshape(-1))
    assert np.array_equal(np.asarray(g_rgb).reshape(-1), np.asarray(c_rgb).reshape(-1))


def test_render_adaptive_lanes8(backend):
    e = cornell().export()
    f = cornell_frame(40, 32, 64, 5, adapt_min=16, adapt_step=8)
    info = _compare_render(backend, e, f, 4242)
    assert info["lanes"] == 8
    _exact_on_backend_tree(backend, e, f, 4242)


def test_render_adaptive_one_lane(backend):
    """adapt_step % 8 != 0 keeps the megakernel at one lane per pixel with the
    adaptive stop (metal.mm:390-394, pathtrace.metal:759-765)."""
    e = cornell().export()
    info = _compare_render(backend, e, cornell_frame(40, 32, 32, 5, adapt_min=16, adapt_step=4), 4243)
    assert info["lanes"] == 1


def test_render_textured_lit_scene_hdr(backend):
    """The lit scene of tests/test_metal.cpp:29-46 with its 2x2 texture (UVs
    outside [0, 1] wrap), compared in linear radiance at the 1e-3 bar."""
    e, f = lit_textured_scene()
    f.spp = 64
    _compare_render(backend, e, f, 12345)


def test_render_hue_shader_scene(backend):
    """shader == 1 triangles in a path-traced frame."""
    from transport_pins import hue_scene
    e, cam = hue_scene(120.0)
    f = frame(48, 48, 16, 4, cam=cam, candidates=4, spatial=0)
    _compare_render(backend, e, f, 99)
    _exact_on_backend_tree(backend, e, f, 99)


def test_render_metal_scene_ris(backend):
    """Mixed materials (metal, smooth and rough glass) with RIS light sampling."""
    e = metal_test_scene().export()
    f = frame(96, 64, 16, 5, candidates=8, spatial=0)
    _compare_render(backend, e, f, 12345)


def test_restir_pipeline_cornell(backend):
    """generate -> spatial -> shade -> temporal -> a-trous x2 -> tonemap, three
    frames with temporal reuse; every intermediate against the oracle."""
    e = cornell().export()
    hist = ob.History()
    backend.set_scene(e)
    for k in range(3):
        cam = scenes.cornell_camera()
        cam.focal = 32 / np.tan(np.radians(20.0))
        cam.x += 0.01 * k
        f = frame(64, 64, 4, 5, cam=cam, candidates=8, spatial=3, radius=8.0)
        seed = 100 + k
        g_rgb, g_hdr = backend.render(f, seed)
        gg, gri, grm, gshade = backend.read_restir(64, 64)
        c_rgb, c_hdr, info = ob.render_frame(e, f, seed, history=hist, debug=True)
        assert info["path"] == 1
        assert backend.stats().has_history == info["has_history"] == (1 if k > 0 else 0)
        kind = info["gbuf"]["kind"]
        assert np.array_equal(gg["kind"], kind)
        on = kind != 0
        for field in ("pos", "normal", "albedo", "wo", "emission", "rough", "metallic", "ior"):
            assert np.array_equal(gg[field][on], info["gbuf"][field][on]), field
        for name, g_res, c_res in (("rinit", gri, info["rinit"]), ("rmerged", grm, info["rmerged"])):
            assert np.array_equal(g_res["valid"], c_res["valid"]), name
            v = c_res["valid"] != 0
            for field in ("light_pos", "light_normal", "emission", "alpha", "p_hat", "sum_ws", "sum_wu", "m"):
                assert np.array_equal(g_res[field][v], c_res[field][v]), (name, field)
        assert rel_err(gshade, info["shade"]).max() <= 1e-3
        check_pass_rays(backend.stats(), info)
        err = rel_err(g_hdr, c_hdr)
        print(f"frame {k}: shade exact {np.mean(gshade == info['shade']):.6f}, final max rel {err.max():.3g}")
        assert err.max() <= 1e-3


def test_restir_pipeline_cornell_same_tree(backend):
    """The same three frames with the oracle traversing the backend's own
    4-wide tree: shading, denoised radiance and 8-bit output bit-identical.
    The first frame starts without history (the backend holds some from the
    test above)."""
    e = cornell().export()
    hist = ob.History()
    backend.set_scene(e)
    nodes, tris, _ = backend.read_tree()
    for k in range(3):
        cam = scenes.cornell_camera()
        cam.focal = 32 / np.tan(np.radians(20.0))
        cam.x += 0.01 * k
        f = frame(64, 64, 4, 5, cam=cam, candidates=8, spatial=3, radius=8.0, temporal=1 if k else 0)
        seed = 100 + k
        g_rgb, g_hdr = backend.render(f, seed)
        gshade = backend.read_restir(64, 64)[3]
        with ob.wide_tree(nodes, tris):
            c_rgb, c_hdr, info = ob.render_frame(e, f, seed, history=hist, debug=True)
        exact = [np.mean(np.asarray(a, np.float32).view(np.uint32).reshape(-1) ==
                         np.ascontiguousarray(b, np.float32).view(np.uint32).reshape(-1))
                 for a, b in ((gshade, info["shade"]), (g_hdr, c_hdr))]
        print(f"frame {k}: shade exact {exact[0]:.6f}, final exact {exact[1]:.6f}")
        assert exact == [1.0, 1.0]
        assert np.array_equal(np.asarray(g_rgb).reshape(-1), np.asarray(c_rgb).reshape(-1))
        check_pass_rays(backend.stats(), info)
