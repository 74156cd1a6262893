"""Checks that pin the oracle's arithmetic independently of its own code:
published constants, closed forms and statistical properties."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle_bindings as ob
from raymotion_amd import rng as np_rng

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_splitmix64_published_first_output():
    # splitmix64 (Vigna) from state 0 outputs 0xE220A8397B1DCDAF first; the
    # stream's 32-bit state is its low half (pathtrace.metal:95).
    _, states = ob.rng_draws(0, 1)
    assert int(states[0]) == 0x7B1DCDAF


def test_pcg_lcg_step_and_lattice():
    _, states = ob.rng_draws(12345, 8)
    s = [int(v) for v in states]
    for a, b in zip(s, s[1:]):
        assert b == (a * 747796405 + 2891336453) % 2**32
    # Hull-Dobell: odd increment, multiplier = 1 mod 4 -> full period 2^32.
    assert 2891336453 % 2 == 1 and 747796405 % 4 == 1
    draws, _ = ob.rng_draws(7, 4096)
    assert np.all(draws >= 0) and np.all(draws < 1)
    assert np.all((draws * 2**24) == np.floor(draws * 2**24))  # 24-bit lattice


def test_rng_uniformity_chi_square():
    draws, _ = ob.rng_draws(99, 1 << 18)
    hist, _ = np.histogram(draws, bins=256, range=(0, 1))
    expected = len(draws) / 256
    chi2 = float(((hist - expected) ** 2 / expected).sum())
    assert chi2 < 256 + 5 * math.sqrt(2 * 256)  # 5 sigma
    # lag-1 serial correlation
    assert abs(np.corrcoef(draws[:-1], draws[1:])[0, 1]) < 0.01


def test_rng_golden_vectors():
    for case in json.loads((GOLDEN / "rng_vectors.json").read_text()):
        draws, states = ob.rng_draws(case["seed"], 16)
        assert int(states[0]) == case["state0"]
        assert [int(v) for v in states[1:]] == case["states"]
        assert [int(v) for v in draws.view(np.uint32)] == case["draws_bits"]


def test_numpy_stream_matches_oracle():
    seeds = np.array([0, 1, 12345, 2**63 + 5], np.uint64)
    got, _ = np_rng.draws(np_rng.splitmix_state(seeds), 10)
    for i, s in enumerate(seeds):
        want, _ = ob.rng_draws(int(s), 10)
        assert np.array_equal(got[i], want)


def test_sincos_within_one_ulp_of_libm():
    x = np.linspace(0.0, 2 * np.pi, 1 << 20, dtype=np.float32)
    x = np.concatenate([x, np.random.default_rng(0).uniform(0, 6.2831855, 1 << 18).astype(np.float32)])
    s, c = ob.sincos(x)
    xs = x.astype(np.float64)
    for got, want in ((s, np.sin(xs)), (c, np.cos(xs))):
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - want) / ulp
        assert err.max() <= 1.0
        # Correctly rounded on all but a vanishing fraction.
        assert np.mean(got == want.astype(np.float32)) > 0.9999


def test_div_pi_correctly_rounded():
    """The BSDF's division by pi (a product with a two-term reciprocal, shared
    by the kernels and the oracle) equals the IEEE quotient x / pi for every
    float with 2^-100 <= |x| < 2^100 (exhaustive, both signs), and for 0,
    infinities and NaN."""
    lo, hi = (127 - 100) << 23, (127 + 100) << 23
    assert ob.div_pi_mismatches(lo, hi) == 0
    assert ob.div_pi_mismatches(0, 1) == 0            # +-0
    assert ob.div_pi_mismatches(0x7F800000, 0x7F800002) == 0  # inf, a NaN


def test_exp_within_one_ulp():
    """The a-trous weights' exponential (single-precision reduction and Taylor
    series, shared by the kernels and the oracle) is within one ulp of libm's
    exp for every float (exhaustive, both signs: underflow to 0 below -104,
    overflow to inf above 88.72, NaN kept), and equal to it for all but a
    small fraction; exp(0) is exactly 1."""
    worst, differ = ob.exp_ulp(0, 0x7F800002)  # every finite float, inf, a NaN
    assert worst <= 1
    assert differ < 0.01 * 2 ** 32, differ
    assert ob.exp_ulp(0, 1) == (0, 0)  # +-0


def test_atan_within_two_ulp():
    """The hue shader's arctangent (single-precision reduction and Taylor
    series, shared by the kernels and the oracle) is within two ulp of libm's
    atan for every float x >= 0 (exhaustive, with inf and a NaN), equal to it
    for all but a small fraction, and rm_atanf(-x) == -rm_atanf(x)."""
    worst, differ = ob.atan_ulp(0, 0x7F800002)
    assert worst <= 2
    assert differ < 0.01 * 2 ** 31, differ


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _bsdf(wi, wo, n, alpha2, metallic, spec_prob, albedo=(0.8, 0.6, 0.4)):
    import ctypes
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    wo = np.ascontiguousarray(np.broadcast_to(wo, wi.shape), np.float32)
    n = np.ascontiguousarray(np.broadcast_to(n, wi.shape), np.float32)
    alb = np.ascontiguousarray(albedo, np.float32)
    out = np.empty((len(wi), 4), np.float32)
    lib = ob.orc()
    lib.orc_eval_bsdf.argtypes = [ctypes.c_size_t] + [ctypes.c_void_p] * 3 + [ctypes.c_float] * 3 + \
        [ctypes.c_void_p] * 2
    lib.orc_eval_bsdf(len(wi), wi.ctypes.data, wo.ctypes.data, n.ctypes.data, alpha2, metallic, spec_prob,
                      alb.ctypes.data, out.ctypes.data)
    return out


def test_fresnel_closed_forms():
    import ctypes
    f = ob.orc().orc_fresnel
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float] * 3
    assert abs(f(1.0, 1.0, 1.5) - 0.04) < 1e-7          # ((1.5-1)/(1.5+1))^2
    assert abs(f(1.0, 1.5, 1.0) - 0.04) < 1e-7
    assert f(0.2, 1.5, 1.0) == 1.0                        # total internal reflection
    assert abs(f(0.0, 1.0, 1.5) - 1.0) < 1e-6             # grazing
    # Monotone in the cosine for an external reflection.
    vals = [f(c, 1.0, 1.5) for c in np.linspace(0.05, 1.0, 50)]
    assert all(a >= b for a, b in zip(vals, vals[1:]))


def test_bsdf_reciprocity_and_hemisphere():
    rng = np.random.default_rng(1)
    n = np.array([0.0, 0.0, 1.0])
    wi = _unit(rng.normal(size=(2000, 3)))
    wi[:, 2] = np.abs(wi[:, 2])
    wo = _unit([0.3, -0.2, 0.9])
    a = _bsdf(wi, wo, n, 0.2, 0.3, 0.335)
    b = np.stack([_bsdf(wo[None], w, n, 0.2, 0.3, 0.335)[0] for w in wi[:200]])
    assert np.allclose(a[:200, :3], b[:, :3], rtol=2e-5, atol=1e-7)  # f(wi,wo) = f(wo,wi)
    below = _bsdf(-wi, wo, n, 0.2, 0.3, 0.335)
    assert np.all(below == 0)


@pytest.mark.parametrize("alpha2,spec", [(1.0, 0.05), (0.0016, 0.5), (0.09, 1.0), (0.25, 0.3)])
def test_bsdf_pdf_mass_closed_form(alpha2, spec):
    # Viewed along the normal, the pdf is axisymmetric; its mass over the upper
    # hemisphere is (1 - p_s) for the cosine lobe plus p_s times the GGX mass
    # that reflects above the horizon, P(theta_h <= 45 deg) = 1 / (1 + alpha^2).
    theta = np.linspace(0.0, np.pi / 2, 400001)
    wi = np.stack([np.sin(theta), np.zeros_like(theta), np.cos(theta)], 1)
    out = _bsdf(wi, [0.0, 0.0, 1.0], [0.0, 0.0, 1.0], alpha2, 0.5, spec)
    f = out[:, 3].astype(np.float64) * 2 * np.pi * np.sin(theta)
    mass = float(np.sum((f[1:] + f[:-1]) * 0.5 * np.diff(theta)))
    want = (1 - spec) + spec / (1 + alpha2)
    assert abs(mass - want) < 2e-3, (mass, want)


def _samples(kind, n, alpha2=0.0):
    import ctypes
    out = np.empty((n, 3), np.float32)
    nrm = np.array([0.0, 0.0, 1.0], np.float32)
    lib = ob.orc()
    lib.orc_sample.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, ctypes.c_float, ctypes.c_void_p,
                               ctypes.c_void_p]
    lib.orc_sample(5, kind, n, alpha2, nrm.ctypes.data, out.ctypes.data)
    return out.astype(np.float64)


def test_cosine_sampler_moments():
    d = _samples(0, 1 << 18)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert abs(d[:, 2].mean() - 2 / 3) < 3e-3          # E[cos] under cos/pi
    assert abs((d[:, 2] ** 2).mean() - 0.5) < 3e-3     # E[cos^2] = 1/2
    assert abs(d[:, 0].mean()) < 5e-3 and abs(d[:, 1].mean()) < 5e-3


@pytest.mark.parametrize("alpha2", [0.01, 0.25])
def test_ggx_half_vector_distribution(alpha2):
    # Sampled cos(theta_h) has CDF P(c^2 <= x) with c^2 = (1-v)/(1+(a2-1)v):
    # v = (1 - x)/(1 + (a2 - 1) x) ... compare quantiles of v recovered from c.
    d = _samples(1, 1 << 17, alpha2)
    c2 = np.clip(d[:, 2], 0, 1) ** 2
    v = (1 - c2) / (1 + (alpha2 - 1) * c2)
    hist, _ = np.histogram(v, bins=32, range=(0, 1))
    assert np.abs(hist / len(v) * 32 - 1).max() < 0.06
