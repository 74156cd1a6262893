"""Scene and frame builders shared by the CPU and GPU tests."""
from __future__ import annotations

import math

import numpy as np

import raymotion_amd as rm
from raymotion_amd import scenes
from raymotion_amd.scene import material_rows


def frame(width, height, spp, bounces=5, *, cam=None, candidates=1, spatial=0, mcap=128, radius=16.0, adapt_min=0,
          adapt_step=8, clamp=16.0, denoise=1, iterations=2, temporal=1, nomis=0):
    """An rm_frame with explicit settings (no environment dependence)."""
    f = rm.Frame()
    f.width, f.height, f.spp, f.bounces = width, height, spp, bounces
    f.resolution = 1.0
    f.clamp_value = clamp
    f.adapt_min, f.adapt_step = adapt_min, adapt_step
    f.adapt_rel, f.adapt_abs = 0.02, 0.002
    f.nomis = nomis
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = candidates, spatial, mcap, radius
    f.denoise, f.denoise_iterations, f.temporal = denoise, iterations, temporal
    f.camera = cam if cam is not None else default_camera(width)
    return f


def default_camera(width, focal=None):
    """Camera at the origin looking down +z (the reference's default Camera),
    focal width * 0.8660254 unless given (tests/test_metal.cpp:21)."""
    c = rm.Camera()
    c.x = c.y = c.z = 0.0
    c.focal = width * 0.8660254 if focal is None else focal
    c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, c.m6, c.m7, c.m8 = 1, 0, 0, 0, 1, 0, 0, 0, 1
    return c


def reference_default_camera():
    """raymotion::Camera{} after update_trig/update_focal (fov 60)."""
    return scenes.camera((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 60.0)


def tri(scene, a, b, c, color, emission=(0, 0, 0), metal=0.0, ior=1.0, rough=0.5, alpha=1.0, shader=0):
    """triangle() of tests/test_metal.cpp:10-14: flat normal, given material."""
    a, b, c = (np.asarray(v, np.float64) for v in (a, b, c))
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    m = material_rows(1, albedo=color, emission=emission, metallic=metal, ior=ior, rough=rough, alpha=alpha,
                      shader=shader)
    scene.add_triangles(np.array([[a, b, c]]), m, normals=np.array([[n, n, n]]))


def quad(scene, a, b, c, d, color, **kw):
    tri(scene, a, b, c, color, **kw)
    tri(scene, a, c, d, color, **kw)


def metal_test_scene():
    """The lit mixed-material scene of tests/test_metal.cpp:29-35 (no texture)."""
    s = rm.Scene()
    quad(s, (-4, -2, 7), (-4, 3, 7), (4, 3, 7), (4, -2, 7), (.7, .7, .7))
    quad(s, (-4, -2, 1), (-4, -2, 7), (4, -2, 7), (4, -2, 1), (.7, .2, .1))
    quad(s, (-1, 2, 3), (1, 2, 3), (1, 2, 5), (-1, 2, 5), (1, 1, 1), emission=(6, 6, 6))
    quad(s, (-2, -1, 4), (-2, 1, 4), (-1, 1, 4), (-1, -1, 4), (.8, .7, .3), metal=1.0)
    quad(s, (0, -1, 4), (0, 1, 4), (1, 1, 4), (1, -1, 4), (1, 1, 1), ior=1.5, rough=.02)
    quad(s, (1, -1, 5), (1, 1, 5), (2, 1, 5), (2, -1, 5), (1, 1, 1), ior=1.5, rough=.3)
    s.build_bvh(2)
    return s


def alpha_surface(scene, z, alpha, emission):
    """surface() of tests/test_alpha.cpp:7-11."""
    n = np.array([0.0, 0.0, -1.0])
    m = material_rows(1, albedo=(0, 0, 0), emission=emission, metallic=0.0, ior=1.0, rough=0.5, alpha=alpha)
    scene.add_triangles(np.array([[(-10, -10, z), (0, 10, z), (10, -10, z)]], np.float64), m,
                        normals=np.array([[n, n, n]]))


def cornell(rough=1.0):
    s = scenes.cornell_box(rough=rough)
    s.build_bvh()
    return s


def cornell_frame(width=256, height=256, spp=4, bounces=5, **kw):
    """C0: pinhole at (0.5, 0.5, -1.4) with a 40 degree vertical field of view."""
    cam = scenes.cornell_camera()
    cam.focal = (height / 2) / math.tan(math.radians(20.0))
    return frame(width, height, spp, bounces, cam=cam, **kw)


def rel_err(a, b, floor=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
