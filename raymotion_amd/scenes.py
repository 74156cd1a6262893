"""Procedural scenes of the benchmark workloads (synthetic, seeded).

cornell_box     36 triangles in the unit cube, Lambert walls, ceiling light
                (17, 12, 4), two boxes rotated +-18 degrees about y.
mixed_spheres   64 icospheres (subdivision 4, 5,120 triangles each) with radii
                U[0.2, 0.6] and centres U[-4, 4]^3, a 20 x 20 ground quad and
                8 emissive 0.4 x 0.4 quads with emission U[5, 20]^3; 40 %
                Lambert, 30 % GGX metal (roughness U[0.05, 0.6]), 30 %
                dielectric (IOR 1.5, roughness 0 or 0.2).
heightfield     an n x n vertex grid displaced by 4-octave value noise, all GGX
                with roughness U[0.1, 0.8], plus 256 emissive triangles with
                log-uniform emission in [1, 100]; vertices animate as
                0.05 sin(2 pi t + x).

Camera convention: the scene API's focal = 240 / tan(fov / 2) pixels,
independent of the image size (src/engine/camera.cpp:3-6).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._capi import Camera, lib
from .scene import Scene, material_rows


def camera(pos, rot=(0.0, 0.0, 0.0), fov=60.0) -> Camera:
    c = Camera()
    p = (ctypes.c_double * 3)(*pos)
    r = (ctypes.c_double * 3)(*rot)
    lib().rm_camera_set(p, r, float(fov), ctypes.byref(c))
    return c


def orbit_camera(center, radius, height, angle_deg, fov=60.0) -> Camera:
    c = Camera()
    ce = (ctypes.c_double * 3)(*center)
    lib().rm_camera_orbit(ce, float(radius), float(height), float(angle_deg), float(fov), ctypes.byref(c))
    return c


def _quad(a, b, c, d):
    """Two triangles (a, b, c) and (a, c, d); their normal is (b - a) x (c - a)."""
    return np.array([[a, b, c], [a, c, d]], np.float64)


def _box(size, angle_deg, base):
    """12 outward-wound triangles of an axis box of `size`, rotated about y by
    `angle_deg` around its centre, standing on `base` (x, y, z of the bottom centre)."""
    sx, sy, sz = size
    x0, x1, y0, y1, z0, z1 = -sx / 2, sx / 2, 0.0, sy, -sz / 2, sz / 2
    p = lambda x, y, z: (x, y, z)
    faces = [
        _quad(p(x0, y0, z0), p(x1, y0, z0), p(x1, y0, z1), p(x0, y0, z1)),  # bottom (-y)
        _quad(p(x0, y1, z0), p(x0, y1, z1), p(x1, y1, z1), p(x1, y1, z0)),  # top (+y)
        _quad(p(x0, y0, z0), p(x0, y1, z0), p(x1, y1, z0), p(x1, y0, z0)),  # front (-z)
        _quad(p(x0, y0, z1), p(x1, y0, z1), p(x1, y1, z1), p(x0, y1, z1)),  # back (+z)
        _quad(p(x0, y0, z0), p(x0, y0, z1), p(x0, y1, z1), p(x0, y1, z0)),  # left (-x)
        _quad(p(x1, y0, z0), p(x1, y1, z0), p(x1, y1, z1), p(x1, y0, z1)),  # right (+x)
    ]
    tris = np.concatenate(faces)
    a = np.deg2rad(angle_deg)
    rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    return tris @ rot.T + np.asarray(base, np.float64)


WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)


def cornell_box(scene: Scene | None = None, rough: float = 1.0) -> Scene:
    """The 36-triangle Cornell box (walls inward-facing, light facing down)."""
    s = scene or Scene()
    walls = [
        (_quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), WHITE),  # floor, normal +y
        (_quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), WHITE),  # ceiling, normal -y
        (_quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), WHITE),  # back, normal -z
        (_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), RED),    # left, normal +x
        (_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), GREEN),  # right, normal -x
    ]
    for tris, color in walls:
        s.add_triangles(tris, material_rows(2, albedo=color, rough=rough))
    ly = 0.999
    light = _quad((0.25, ly, 0.25), (0.75, ly, 0.25), (0.75, ly, 0.75), (0.25, ly, 0.75))
    s.add_triangles(light, material_rows(2, albedo=WHITE, emission=(17.0, 12.0, 4.0), rough=rough))
    s.add_triangles(_box((0.3, 0.3, 0.3), -18.0, (0.65, 0.0, 0.35)), material_rows(12, albedo=WHITE, rough=rough))
    s.add_triangles(_box((0.3, 0.6, 0.3), 18.0, (0.33, 0.0, 0.62)), material_rows(12, albedo=WHITE, rough=rough))
    return s


def cornell_camera() -> Camera:
    return camera((0.5, 0.5, -1.4), (0.0, 0.0, 0.0), 40.0)


def icosphere(subdivisions: int = 4) -> np.ndarray:
    """Unit icosphere triangles (F, 3, 3), outward winding; F = 20 * 4^subdivisions."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    tris = v[f]
    for _ in range(subdivisions):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tris = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                               np.stack([ab, bc, ca], 1)])
    # Outward winding: (v1 - v0) x (v2 - v0) along the centroid direction.
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1)) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris


def mixed_sphere_parts(seed: int = 0, n_spheres: int = 64, subdivisions: int = 4):
    """The mixed scene as a list of (positions, materials, normals or None) parts."""
    rng = np.random.default_rng(seed)
    unit = icosphere(subdivisions)
    parts = []
    for _ in range(n_spheres):
        r = rng.uniform(0.2, 0.6)
        c = rng.uniform(-4.0, 4.0, 3)
        kind = rng.uniform()
        if kind < 0.4:
            mat = material_rows(len(unit), albedo=rng.uniform(0.1, 0.9, 3), rough=1.0)
        elif kind < 0.7:
            mat = material_rows(len(unit), albedo=rng.uniform(0.5, 1.0, 3), metallic=1.0,
                                rough=rng.uniform(0.05, 0.6))
        else:
            mat = material_rows(len(unit), albedo=(1.0, 1.0, 1.0), ior=1.5, rough=float(rng.choice([0.0, 0.2])))
        parts.append((c + r * unit, mat, unit))
    g = -4.6
    parts.append((_quad((-10, g, -10), (-10, g, 10), (10, g, 10), (10, g, -10)),
                  material_rows(2, albedo=(0.5, 0.5, 0.5), rough=1.0), None))
    for _ in range(8):
        x, z = rng.uniform(-4.0, 4.0, 2)
        y, h = 4.9, 0.2
        quad = _quad((x - h, y, z - h), (x + h, y, z - h), (x + h, y, z + h), (x - h, y, z + h))
        parts.append((quad, material_rows(2, albedo=(0.8, 0.8, 0.8), emission=rng.uniform(5.0, 20.0, 3), rough=1.0),
                      None))
    return parts


def mixed_spheres(seed: int = 0, n_spheres: int = 64, subdivisions: int = 4, scene: Scene | None = None) -> Scene:
    s = scene or Scene()
    for pos, mat, nrm in mixed_sphere_parts(seed, n_spheres, subdivisions):
        s.add_triangles(pos, mat, normals=nrm)
    return s


def inverse_scene(seed: int = 0, n_spheres: int = 64, subdivisions: int = 4):
    """The mixed-scene generator reduced to 16 materials, for inverse rendering:
    13 sphere materials drawn like the mixed scene's (40% Lambert, 30% GGX
    metal, 30% dielectric of roughness 0 or 0.2), the ground (group 13) and
    the 8 emissive quads in two groups of 4 (14, 15) with one emission each.
    Lambert roughness is drawn in [0.4, 0.9] and metal albedo in [0.5, 0.95] so
    that +-30% perturbations stay inside the parameter box. Returns
    (scene with its BVH built, per-triangle group)."""
    rng = np.random.default_rng(seed)
    unit = icosphere(subdivisions)
    mats = []
    for _ in range(13):
        kind = rng.uniform()
        if kind < 0.4:
            mats.append(material_rows(1, albedo=rng.uniform(0.1, 0.9, 3), rough=rng.uniform(0.4, 0.9)))
        elif kind < 0.7:
            mats.append(material_rows(1, albedo=rng.uniform(0.5, 0.95, 3), metallic=1.0, rough=rng.uniform(0.05, 0.6)))
        else:
            mats.append(material_rows(1, albedo=(1.0, 1.0, 1.0), ior=1.5, rough=float(rng.choice([0.0, 0.2]))))
    s = Scene()
    groups = []
    for _ in range(n_spheres):
        r = rng.uniform(0.2, 0.6)
        c = rng.uniform(-4.0, 4.0, 3)
        m = len(groups) if len(groups) < 13 else int(rng.integers(13))  # every material used once first
        idx = s.add_triangles(c + r * unit, np.repeat(mats[m], len(unit), 0), normals=unit)
        groups.append(np.full(len(idx), m, np.int32))
    g = -4.6
    s.add_triangles(_quad((-10, g, -10), (-10, g, 10), (10, g, 10), (10, g, -10)),
                    material_rows(2, albedo=(0.5, 0.5, 0.5), rough=0.8))
    groups.append(np.full(2, 13, np.int32))
    emission = [rng.uniform(5.0, 20.0, 3), rng.uniform(5.0, 20.0, 3)]
    for k in range(8):
        x, z = rng.uniform(-4.0, 4.0, 2)
        y, h = 4.9, 0.2
        quad = _quad((x - h, y, z - h), (x + h, y, z - h), (x + h, y, z + h), (x - h, y, z + h))
        s.add_triangles(quad, material_rows(2, albedo=(0.8, 0.8, 0.8), emission=emission[k // 4], rough=1.0))
        groups.append(np.full(2, 14 + k // 4, np.int32))
    s.build_bvh()
    return s, np.concatenate(groups)


def mixed_camera(frame: int = 0, frames: int = 34) -> Camera:
    """Orbit at radius 10, height 3 around the origin."""
    return orbit_camera((0.0, 0.0, 0.0), 10.0, 3.0, 360.0 * frame / frames, 45.0)


def _value_noise(x, y, octaves, rng):
    """Value noise on a unit lattice, smoothstep-interpolated, summed over octaves."""
    total = np.zeros_like(x)
    amp, freq = 1.0, 1.0
    for _ in range(octaves):
        lat = rng.uniform(-1.0, 1.0, (66, 66))
        fx, fy = x * freq, y * freq
        ix, iy = np.floor(fx).astype(int) % 64, np.floor(fy).astype(int) % 64
        tx, ty = fx - np.floor(fx), fy - np.floor(fy)
        sx, sy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        v00, v10 = lat[ix, iy], lat[ix + 1, iy]
        v01, v11 = lat[ix, iy + 1], lat[ix + 1, iy + 1]
        total += amp * ((v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy)
        amp *= 0.5
        freq *= 2.0
    return total


def heightfield_positions(n: int = 1024, seed: int = 0, t: float = 0.0, extent: float = 16.0) -> np.ndarray:
    """Grid triangles (2 (n-1)^2, 3, 3) of the displaced, animated height field."""
    rng = np.random.default_rng(seed)
    lin = np.linspace(-extent / 2, extent / 2, n)
    xs, zs = np.meshgrid(lin, lin, indexing="ij")
    h = 1.5 * _value_noise((xs + extent) * 0.25, (zs + extent) * 0.25, 4, rng)
    h = h + 0.05 * np.sin(2.0 * np.pi * t + xs)
    p = np.stack([xs, h, zs], axis=-1)
    a, b, c, d = p[:-1, :-1], p[1:, :-1], p[1:, 1:], p[:-1, 1:]
    t1 = np.stack([a, d, c], axis=-2).reshape(-1, 3, 3)
    t2 = np.stack([a, c, b], axis=-2).reshape(-1, 3, 3)
    return np.concatenate([t1, t2])


def _heightfield_extras(n: int, seed: int):
    """Per-triangle grid roughness, and the 256 emissive triangles with their
    material rows (emission log-uniform over [1, 100])."""
    rng = np.random.default_rng(seed + 1)
    rough = rng.uniform(0.1, 0.8, 2 * (n - 1) * (n - 1))
    k = 256
    centers = np.stack([rng.uniform(-7, 7, k), rng.uniform(3.0, 5.0, k), rng.uniform(-7, 7, k)], 1)
    size = 0.3
    tri = np.stack([centers + [-size, 0, -size], centers + [size, 0, -size], centers + [0, 0, size]], 1)
    emis = np.exp(rng.uniform(np.log(1.0), np.log(100.0), (k, 3)))
    lm = material_rows(k, albedo=(0.8, 0.8, 0.8), rough=1.0)
    lm[:, 3:6] = emis
    return rough, tri, lm


def heightfield(n: int = 1024, seed: int = 0, t: float = 0.0, scene: Scene | None = None) -> Scene:
    """n x n height field, all GGX metal with roughness U[0.1, 0.8], lit by 256
    emissive triangles; the animated displacement is at time t."""
    s = scene or Scene()
    grid = heightfield_positions(n, seed, t)
    rough, light_tri, light_mat = _heightfield_extras(n, seed)
    mat = material_rows(len(grid), albedo=(0.7, 0.7, 0.7), metallic=1.0)
    mat[:, 8] = rough
    s.add_triangles(grid, mat)
    s.add_triangles(light_tri, light_mat)
    return s


class HeightFieldAnimation:
    """Triangle positions of heightfield(n, seed) at time t (grid, then the
    static lights), with the noise evaluated once. The grid has n distinct x
    coordinates, so the animated term is evaluated per column and gathered."""

    def __init__(self, n: int = 1024, seed: int = 0):
        self.base = heightfield_positions(n, seed, 0.0)
        x = self.base[:, :, 0]
        self.h0 = self.base[:, :, 1] - 0.05 * np.sin(x)
        self.ux, self.xi = np.unique(x, return_inverse=True)
        self.xi = self.xi.reshape(x.shape)
        self.lights = _heightfield_extras(n, seed)[1]
        self.out = np.concatenate([self.base, self.lights])

    def positions(self, t: float) -> np.ndarray:
        """Positions at time t, in a buffer that the next call overwrites."""
        y = self.out[:len(self.base), :, 1]
        np.multiply(np.sin(2.0 * np.pi * t + self.ux)[self.xi], 0.05, out=y)
        y += self.h0
        return self.out


def heightfield_camera(frame: int = 0) -> Camera:
    return orbit_camera((0.0, 0.0, 0.0), 11.0, 6.0, 30.0 + 2.0 * frame, 50.0)
