"""Known-answer cases of the reference's own GPU tests, as functions of a
`render(exported, frame, seed) -> rgb8` callable so that the same checks run
against the CPU oracle and against the HIP backend.

  metal_*  tests/test_metal.cpp:19-76
  alpha_*  tests/test_alpha.cpp:12-29
"""
from __future__ import annotations

import numpy as np

import raymotion_amd as rm
from helpers import alpha_surface, default_camera, frame, metal_test_scene, quad


def _cfg(width, height, spp, adapt_min=16, cam=None):
    # RenderConfig defaults (pathtrace.hpp:54-81): candidates 8, spatial 4,
    # radius 16, mcap 128, clamp 16, adaptive 16/8; denoise and temporal on.
    return frame(width, height, spp, 5, cam=cam, candidates=8, spatial=4, radius=16.0, mcap=128,
                 adapt_min=adapt_min, adapt_step=8)


def _exported(scene):
    return scene.export()


def metal_empty(render):
    """test_metal.cpp:20-28: an empty scene renders black."""
    f = _cfg(96, 64, 512, adapt_min=0)
    img = render(_exported(rm.Scene()), f, 12345)
    assert np.all(img == 0), "empty scene is not black"


def lit_textured_scene():
    """The lit mixed-material scene of test_metal.cpp:29-41 with its 2x2
    texture on every non-emissive triangle, and the test's frame settings."""
    f = _cfg(96, 64, 512, adapt_min=0)
    e = _exported(metal_test_scene())
    emissive = (e.tris["er"] + e.tris["eg"] + e.tris["eb"]) == 0
    e.tris["texture_offset"][emissive] = 0
    e.tris["texture_width"][emissive] = 2
    e.tris["texture_height"][emissive] = 2
    for name, value in (("tu0", -.2), ("tv0", .1), ("tu1", 1.4), ("tv1", .2), ("tu2", .3), ("tv2", 1.7)):
        e.tris[name][emissive] = value
    e.texels = np.array([1, .1, .1, .1, 1, .1, .1, .1, 1, .5, .5, .5], np.float32)
    return e, f


def metal_lit_deterministic(render):
    """test_metal.cpp:29-46: the lit mixed-material scene with a 2x2 texture
    (UVs outside [0,1] wrap) is not black and renders identically twice."""
    e, f = lit_textured_scene()
    a = render(e, f, 12345)
    assert np.any(a > 0), "lit scene is black"
    b = render(e, f, 12345)
    assert np.array_equal(a, b), "output is not deterministic"
    return a


def _fullscreen_quad(x=0.0):
    s = rm.Scene()
    quad(s, (x - 20, -20, 3), (x - 20, 20, 3), (x + 20, 20, 3), (x + 20, -20, 3), (1, 1, 1), emission=(1, 0, 0))
    s.build_bvh()
    return s


def metal_emission_exact(render):
    """test_metal.cpp:47-52: a full-screen emissive quad gives exactly (127,0,0)
    with adaptive sampling, after a scene re-upload."""
    f = _cfg(96, 64, 512, adapt_min=16)
    img = render(_exported(_fullscreen_quad()), f, 98765).reshape(-1, 3)
    assert np.all(img[:, 0] == 127) and np.all(img[:, 1:] == 0), "emission/adaptive/upload mismatch"


def metal_alpha_coverage(render):
    """test_metal.cpp:53-66: coverage 0, 0.5, 1 on the emissive quad at 2048 spp."""
    f = _cfg(16, 16, 2048, adapt_min=0, cam=default_camera(16))
    e = _exported(_fullscreen_quad())
    for opacity in (0.0, 0.5, 1.0):
        e.tris["alpha"][:] = opacity
        img = render(e, f, 98765).reshape(-1, 3)
        if opacity == 0.5:
            assert abs(img[:, 0].mean() - 85) < 3, f"alpha coverage mismatch: {img[:, 0].mean()}"
        if opacity == 0.0:
            assert np.all(img == 0), "transparent emission visible"
        if opacity == 1.0:
            assert img[0, 0] == 127, "opaque emission changed"


def metal_frames_refit(render):
    """test_metal.cpp:67-76: 34 frames at an odd size alternate the quad in and
    out of view (buffer reuse, dispatch tails, refits and the periodic rebuild)."""
    f = _cfg(37, 23, 9, adapt_min=0)
    for k in range(34):
        x = (k % 2) * 100.0
        img = render(_exported(_fullscreen_quad(x)), f, 12345).reshape(-1, 3)
        want = 0 if k % 2 else 127
        assert np.all(img[:, 0] == want) and np.all(img[:, 1:] == 0), f"refit or tail mismatch at frame {k}"


def _alpha_cfg(w, h, spp):
    cam = rm.Camera()
    cam.focal = 201.3839114815688  # default-constructed Camera (pathtrace.hpp:17)
    cam.m0, cam.m4, cam.m8 = 1.0, 1.0, 1.0
    return _cfg(w, h, spp, adapt_min=0, cam=cam)


def alpha_two_layers(render):
    """test_alpha.cpp:13-22: red surface of opacity o in front of opaque green;
    mean red 255 o/(1+o), mean green 255 (1-o)/(2-o), within 3."""
    f = _alpha_cfg(8, 8, 4096)
    for o in (0.0, 0.5, 1.0):
        s = rm.Scene()
        alpha_surface(s, 1, o, (1, 0, 0))
        alpha_surface(s, 2, 1, (0, 1, 0))
        s.build_bvh()
        img = render(_exported(s), f, 12345).reshape(-1, 3).astype(np.float64)
        assert abs(img[:, 0].mean() - 255 * o / (1 + o)) < 3, (o, img[:, 0].mean())
        assert abs(img[:, 1].mean() - 255 * (1 - o) / (2 - o)) < 3, (o, img[:, 1].mean())


def alpha_transparent_stack(render):
    """test_alpha.cpp:23-29: 64 fully transparent layers over opaque green give
    exactly (0, 127) at one sample: crossings consume no bounces."""
    s = rm.Scene()
    for i in range(64):
        alpha_surface(s, 1 + i * .01, 0, (1, 0, 0))
    alpha_surface(s, 2, 1, (0, 1, 0))
    s.build_bvh()
    img = render(_exported(s), _alpha_cfg(1, 1, 1), 12345)
    assert img[0] == 0 and img[1] == 127, img


ALL = [metal_empty, metal_lit_deterministic, metal_emission_exact, metal_alpha_coverage, metal_frames_refit,
       alpha_two_layers, alpha_transparent_stack]
