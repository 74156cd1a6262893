"""Full-size rendering configs against the oracle, frame by frame with the
temporal state carried as the boundary carries it (render_image_metal's
history rules, metal.mm:232-254):

* the mixed-material scene (327,698 triangles, tree built on the GPU) at
  1280x720, 16 spp, depth 8, ReSTIR 32 candidates / 3 spatial neighbours in
  30 px with temporal reuse, adaptive sampling and 5 a-trous passes, orbit
  frames 0, 1, 20 and 21 (all but the first with history);
* the animated height field (2,093,314 triangles, tree built on the GPU and
  refitted after the vertices move) at 1920x1080, 8 spp, depth 6, ReSTIR 8 / 4
  in 16 px and 2 a-trous passes, before and after one animation step.

The oracle traverses the backend's own 4-wide tree (oracle_bindings.wide_tree):
the G-buffer, the initial and merged reservoirs, the denoised FP32 radiance
and the 8-bit output equal the oracle's bit for bit, and so do the ray,
node-visit and triangle-test counts."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import frame, rel_err
from raymotion_amd import scenes
from raymotion_amd.denoise import restir_frame

pytestmark = pytest.mark.gpu


def _fields_equal(a, c):
    return all(np.array_equal(np.ascontiguousarray(a[k]).view(np.uint32), np.ascontiguousarray(c[k]).view(np.uint32))
               for k in a.dtype.names)


def _check(backend, exported, f, seed, hist, expect_history):
    g_rgb, g_hdr = backend.render(f, seed)
    st = backend.stats()
    g_gbuf, g_rinit, g_rmerged, _ = backend.read_restir(f.width, f.height)
    # the oracle traverses the backend's own tree (same boxes, same slab arithmetic)
    nodes, tris, _ = backend.read_tree()
    with ob.wide_tree(nodes, tris):
        c_rgb, c_hdr, info = ob.render_frame(exported, f, seed, history=hist, debug=True)
    # G-buffer, initial and merged reservoirs: bit for bit
    assert _fields_equal(g_gbuf, info["gbuf"])
    assert _fields_equal(g_rinit, info["rinit"])
    assert _fields_equal(g_rmerged, info["rmerged"])
    assert (st.node_visits[5], st.tri_tests[5]) == (info["node_visits"], info["tri_tests"])
    err = rel_err(g_hdr, c_hdr)
    same = (np.ascontiguousarray(g_hdr, np.float32).view(np.uint32).reshape(-1) ==
            np.ascontiguousarray(c_hdr, np.float32).view(np.uint32).reshape(-1))
    print(f"{f.width}x{f.height} {f.spp} spp seed {seed}: path {info['path']}, history {info['has_history']}, "
          f"{info['rays']} oracle rays, max rel {err.max():.3g}, exact {np.mean(same):.6f}", flush=True)
    assert info["path"] == 1
    assert info["has_history"] == expect_history
    assert same.all()
    assert np.array_equal(np.asarray(g_rgb).reshape(-1), np.asarray(c_rgb).reshape(-1))
    gpu = {"megakernel": st.rays[0], "generate": st.rays[1], "spatial": st.rays[2], "shade": st.rays[3]}
    print(f"  rays per pass: GPU {gpu}, oracle {info['pass_rays']}", flush=True)
    assert gpu == info["pass_rays"]


def test_mixed_scene_720p_frames_match_oracle(backend):
    s = scenes.mixed_spheres(seed=0)
    s.build_bvh()
    e = s.export()
    backend.set_scene(e, gpu_tree=True, rebuild=True)
    hist = ob.History()
    for i, k in enumerate((0, 1, 20, 21)):
        # the first frame starts without history (the backend may hold some
        # from an earlier render of this scene at this size); the later ones
        # carry it, across the jump from orbit frame 1 to 20 as well
        f = restir_frame(1280, 720, 16, scenes.mixed_camera(k), denoise=1, temporal=1 if i else 0)
        _check(backend, e, f, 12345 + k, hist, 1 if i else 0)


def test_heightfield_1080p_frames_match_oracle(backend):
    s = scenes.heightfield(1024, seed=0)
    anim = scenes.HeightFieldAnimation(1024, seed=0)
    s.set_positions(anim.positions(0.0))
    s.build_bvh()
    backend.set_scene(s, gpu_tree=True, rebuild=True)
    hist = ob.History()
    cam = scenes.heightfield_camera(0)
    f = frame(1920, 1080, 8, 6, cam=cam, candidates=8, spatial=4, mcap=128, radius=16.0, adapt_min=16,
              iterations=2, temporal=0)
    _check(backend, s.export(), f, 12345, hist, 0)
    f.temporal = 1
    # one animation step: the host tree is refitted for the oracle, the GPU
    # refits its own; a changed scene has no history (metal.mm:249-254)
    s.set_positions(anim.positions(1.0 / 60.0))
    s.refit_bvh()
    backend.set_scene(s, gpu_tree=True)
    assert backend.stats().refit == 1
    f.camera = scenes.heightfield_camera(1)
    _check(backend, s.export(), f, 12346, hist, 0)
