"""Pins of the oracle's path-replay adjoint (CPU, FP64).

The adjoint holds sampling, MIS weights and roulette fixed and differentiates
BSDF values and emission. Where the rendered estimate of a fixed seed is a
smooth function of a parameter, its central difference must then equal the
adjoint to FD accuracy:
  * albedo at depth 2 with MIS off: sampling does not depend on albedo and the
    only roulette decision falls after the last vertex;
  * roughness at depth 1 with MIS off: it enters only through the BSDF value
    toward the light sample;
  * emission scale of the only light group: the estimate is linear in it
    (pdfs and weights are scale-invariant), so dL/ds = L / s exactly.
Steps are powers of two from float32 values, so the parameter really moves by
2h in the float32 records. The BSDF derivatives are pinned separately against
FP64 central differences of the BSDF itself.
"""
import math

import numpy as np
import pytest

import oracle_bindings as ob
from helpers import cornell
from raymotion_amd import scenes
from raymotion_amd.diff import DiffScene, path_frame

# Cornell box triangle groups: white walls, red, green, light, short box, tall box.
CORNELL_GROUPS = np.array([0] * 6 + [1] * 2 + [2] * 2 + [3] * 2 + [4] * 12 + [5] * 12, np.int32)


def _setup(bounces, nomis, clamp=0.0, size=48, spp=4):
    ds = DiffScene(cornell(rough=0.5).export(), CORNELL_GROUPS)
    f = path_frame(size, size, spp, bounces, scenes.cornell_camera(), clamp=clamp)
    f.nomis = nomis
    f.camera.focal = (size / 2) / math.tan(math.radians(20))
    d = np.random.default_rng(0).uniform(0, 1, size * size * 3).astype(np.float32)
    return ds, f, d


def _loss(ds, f, d, p):
    return float(np.sum(d.astype(np.float64) * ob.render_reference_f64(ds.exported(p), f, 777)))


def _fd(ds, f, d, gi, k, h):
    p0 = ds.initial.copy()
    hi, lo = p0.copy(), p0.copy()
    hi[gi, k] += h
    lo[gi, k] -= h
    return (_loss(ds, f, d, hi) - _loss(ds, f, d, lo)) / (2 * h)


def _adj(ds, f, d, precision=64):
    return ob.render_adjoint(ds.exported(ds.initial), f, 777, d, ds.tri_group, ds.group_emission,
                             precision=precision)


def test_bsdf_derivatives_match_fd():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = np.array([0.0, 0.0, 1.0])
        wi, wo = rng.normal(size=3), rng.normal(size=3)
        wi[2], wo[2] = abs(wi[2]) + 0.05, abs(wo[2]) + 0.05
        wi /= np.linalg.norm(wi)
        wo /= np.linalg.norm(wo)
        a2, m = rng.uniform(0.01, 0.9), rng.choice([0.0, 0.3, 1.0])
        alb = rng.uniform(0.05, 0.95, 3)
        _, da, dl = ob.bsdf_grad_f64(wi, wo, n, a2, m, alb)
        h = 1e-6
        vp, _, _ = ob.bsdf_grad_f64(wi, wo, n, a2 + h, m, alb)
        vm, _, _ = ob.bsdf_grad_f64(wi, wo, n, a2 - h, m, alb)
        fd_l = (vp - vm) / (2 * h)
        np.testing.assert_allclose(dl, fd_l, rtol=1e-5, atol=1e-7)
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            vp, _, _ = ob.bsdf_grad_f64(wi, wo, n, a2, m, alb + e)
            vm, _, _ = ob.bsdf_grad_f64(wi, wo, n, a2, m, alb - e)
            assert abs(da[c] - (vp[c] - vm[c]) / (2 * h)) <= 1e-6 * max(1.0, abs(da[c]))


@pytest.mark.parametrize("clamp", [0.0, 16.0])
def test_albedo_gradient_exact(clamp):
    ds, f, d = _setup(bounces=2, nomis=1, clamp=clamp)
    g = _adj(ds, f, d)
    for gi, k in [(0, 0), (0, 2), (1, 0), (2, 1), (4, 2), (5, 0)]:
        fd = _fd(ds, f, d, gi, k, 2.0 ** -14)
        assert abs(g[gi, k] - fd) <= 1e-8 * abs(fd), (gi, k, g[gi, k], fd)


def test_roughness_gradient_through_nee_exact():
    ds, f, d = _setup(bounces=1, nomis=1)
    g = _adj(ds, f, d)
    for gi in (0, 1, 4, 5):
        fd = _fd(ds, f, d, gi, 3, 2.0 ** -14)
        assert abs(g[gi, 3] - fd) <= 1e-5 * abs(fd), (gi, g[gi, 3], fd)


def test_emission_scale_closed_form():
    ds, f, d = _setup(bounces=4, nomis=0)
    g = _adj(ds, f, d)
    loss = _loss(ds, f, d, ds.initial)
    assert abs(g[3, 4] - loss) <= 1e-9 * loss  # dL/ds = L / s at s = 1
    assert np.all(g[[0, 1, 2, 4, 5], 4] == 0.0)  # no emission elsewhere


def test_fp32_adjoint_follows_fp64():
    ds, f, d = _setup(bounces=5, nomis=0, clamp=16.0)
    g64 = _adj(ds, f, d, 64)
    g32 = _adj(ds, f, d, 32)
    scale = np.abs(g64).max()
    assert np.abs(g32 - g64).max() <= 1e-4 * scale


def test_unused_parameters_have_zero_gradient():
    ds, f, d = _setup(bounces=3, nomis=0)
    g = _adj(ds, f, d)
    # Lambert walls: no emission gradient; the light's albedo only matters for
    # light reflected off it.
    assert np.all(g[[0, 1, 2, 4, 5], 4] == 0)
    assert np.all(np.isfinite(g))


def test_adjoint_rejects_restir_frames():
    ds, f, d = _setup(bounces=2, nomis=0)
    f.restir_candidates = 8
    with pytest.raises(RuntimeError):
        _adj(ds, f, d)
