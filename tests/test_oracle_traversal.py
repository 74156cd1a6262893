"""Pins the oracle's traversal (pathtrace.metal:119-163 restated, with the
closest-distance-then-smallest-index tie-break) on the CPU, independently of
any GPU:

* against brute force over every triangle (no tree at all), bit for bit in
  triangle id and distance, for closest hits and for any-hit queries with a
  distance bound;
* against an FP64 ray/triangle intersection written here in numpy from the
  Moller-Trumbore formulas (Moller & Trumbore 1997), on the rays whose closest
  FP64 hit is unambiguous (clear of triangle edges and of the next-closest
  hit), so that a dropped term, a swapped barycentric or a wrong sign in the
  restated test fails.
"""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import cornell
from raymotion_amd import scenes
from raymotion_amd.rng import random_rays


def _vertices(exported):
    t = exported.tris
    return [np.asarray(t[k], np.float64) for k in ("v0", "v1", "v2")]


def _fp64_closest(exported, o, d):
    """Closest hit per ray in FP64 over all triangles; also the margin data
    that says whether the answer is robust to FP32 rounding."""
    v0, v1, v2 = _vertices(exported)
    o = o.astype(np.float64)[:, None, :]
    d = d.astype(np.float64)[:, None, :]
    e1 = (v1 - v0)[None]
    e2 = (v2 - v0)[None]
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=-1)
    ok = np.abs(det) > 1e-9
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    tv = o - v0[None]
    u = np.sum(tv * p, axis=-1) * inv
    q = np.cross(tv, e1)
    v = np.sum(d * q, axis=-1) * inv
    t = np.sum(e2 * q, axis=-1) * inv
    hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 1e-8)
    tt = np.where(hit, t, np.inf)
    order = np.argsort(tt, axis=1)
    best = order[:, 0]
    rows = np.arange(len(best))
    t_best = tt[rows, best]
    t_next = tt[rows, order[:, 1]]
    # distance from the chosen triangle's edges in barycentric units, and
    # from the edges of every other triangle the ray passes near
    edge = np.minimum(np.minimum(u, v), 1 - u - v)
    near_edge = np.any((np.abs(edge) < 1e-4) & ok & (t > 0), axis=1)
    robust = np.isfinite(t_best) & (t_next > t_best * (1 + 1e-4) + 1e-6) & ~near_edge & (t_best > 1e-3)
    miss_robust = ~np.isfinite(t_best) & ~near_edge
    return np.where(np.isfinite(t_best), best, -1), t_best, robust, miss_robust


def _soup_scene():
    s = scenes.mixed_spheres(seed=5, n_spheres=3, subdivisions=3)
    s.build_bvh(2)
    return s.export()


@pytest.mark.parametrize("which", ["cornell", "soup"])
def test_oracle_closest_hit_equals_brute_force(which):
    if which == "cornell":
        e = cornell().export()
        o, d = random_rays(1 << 18, 0)
    else:
        e = _soup_scene()
        o, d = random_rays(1 << 14, 9, lo=(-5, -5, -5), hi=(5, 5, 5))
    tri, t, _, _, stats = ob.trace(e, o, d)
    b_tri, b_t = ob.trace_brute(e, o, d)
    assert np.array_equal(tri, b_tri)
    hit = b_tri >= 0
    assert np.array_equal(t[hit], b_t[hit])
    assert 0.2 < hit.mean() < 1.0
    # the tree prunes: fewer triangle tests than brute force
    assert stats[1] < len(o) * len(e.tris)


def test_oracle_any_hit_equals_bounded_brute_force():
    e = cornell().export()
    o, d = random_rays(1 << 17, 7)
    tmax = np.random.default_rng(1).uniform(0.01, 1.5, len(o)).astype(np.float32)
    tri, _, _, _, _ = ob.trace(e, o, d, tmax, any_hit=True)
    b_tri, b_t = ob.trace_brute(e, o, d)
    expect = (b_tri >= 0) & (b_t < tmax)
    assert np.array_equal(tri >= 0, expect)
    assert 0.1 < expect.mean() < 0.9


@pytest.mark.parametrize("which", ["cornell", "soup"])
def test_oracle_hits_match_fp64_moller_trumbore(which):
    if which == "cornell":
        e = cornell().export()
        o, d = random_rays(1 << 13, 21)
    else:
        e = _soup_scene()
        o, d = random_rays(1 << 11, 23, lo=(-5, -5, -5), hi=(5, 5, 5))
    tri, t, u, v, _ = ob.trace(e, o, d)
    x_tri, x_t, robust, miss_robust = _fp64_closest(e, o, d)
    assert robust.mean() > 0.3
    assert np.array_equal(tri[robust], x_tri[robust])
    assert np.all(tri[miss_robust] == -1)
    rel = np.abs(t[robust] - x_t[robust]) / x_t[robust]
    assert rel.max() < 1e-5
    # barycentrics reported with the hit reproduce the hit point
    v0, v1, v2 = _vertices(e)
    k = tri[robust]
    p_bary = (1 - u[robust] - v[robust])[:, None] * v0[k] + u[robust][:, None] * v1[k] + v[robust][:, None] * v2[k]
    p_ray = o[robust].astype(np.float64) + t[robust][:, None] * d[robust].astype(np.float64)
    assert np.abs(p_bary - p_ray).max() < 1e-4
