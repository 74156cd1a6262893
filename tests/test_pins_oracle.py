"""Pins of the oracle's light transport, denoiser and ReSTIR merge against
expectations the mathematics fixes (tests/transport_pins.py): the FP32 parity
oracle (orc_render_frame) and, where it covers the path, the FP64 renderer.
The same cases run on the HIP backend in tests/test_gpu_pins.py."""
import numpy as np
import pytest

import oracle_bindings as ob
import pin_cases as pc
import transport_pins as tp


class OracleRunner:
    scale = 1
    batches = 16

    def __init__(self, f64=False):
        self.f64 = f64

    def hdr_batches(self, exported, frame, seeds):
        out = []
        for s in seeds:
            if self.f64:
                out.append(ob.render_reference_f64(exported, frame, s).reshape(-1, 3))
            else:
                _, hdr, _ = ob.render_frame(exported, frame, s)
                out.append(hdr.reshape(-1, 3))
        return np.stack(out)

    def gbuffer(self, exported, frame, seed):
        _, _, info = ob.render_frame(exported, frame, seed, debug=True)
        assert info["path"] == 1
        return info["gbuf"]

    def merge(self, exported, frame, seed, gbuf, rinit, hist_gbuf, hist_res, prev_camera):
        return ob.restir_merge(exported, frame, seed, gbuf, rinit, hist_gbuf, hist_res, prev_camera)

    def denoise(self, frame, colour, gbuf, hist_colour, hist_gbuf, prev_camera):
        return ob.denoise_image(frame, colour, gbuf, hist_colour, hist_gbuf, prev_camera)


ORC = OracleRunner()
ORC64 = OracleRunner(f64=True)


def test_quadrature_matches_polygon_form_factor():
    # The direct-light quadrature of cos cos / r^2 over the emitter equals
    # Lambert's contour formula at points all over the plane.
    sc = pc.direct_light()
    y, wa, _ = sc.light_points(48)
    c, e1, e2 = sc.corner, sc.e1, sc.e2
    verts = [c, c + e1, c + e1 + e2, c + e2]
    for x in ([0.0, 0.0, 0.0], [1.3, -0.7, 0.0], [-2.0, 2.5, 0.0], [0.3, 0.1, 0.0]):
        x = np.array(x)
        v = y - x
        r = np.linalg.norm(v, axis=1)
        cs = v[:, 2] / r
        quad = float(np.sum(wa * cs * cs / (r * r)))
        closed = tp.polygon_form_factor_integral(x, np.array([0.0, 0.0, 1.0]), verts)
        assert abs(quad - closed) <= 1e-10 * closed, (x, quad, closed)


def test_fresnel_slab_formula_limits():
    # normal incidence ((n-1)/(n+1))^2; grazing 1; Stokes: F from either side
    # of one interface is the same at corresponding angles.
    assert abs(float(tp.fresnel_unpolarised(np.array(1.0), 1.0, 1.5)) - 0.04) < 1e-12
    assert abs(float(tp.fresnel_unpolarised(np.array(1e-9), 1.0, 1.5)) - 1.0) < 1e-6
    ci = np.linspace(0.05, 1.0, 40)
    ct = np.sqrt(1 - (np.sqrt(1 - ci ** 2) / 1.5) ** 2)
    assert np.allclose(tp.fresnel_unpolarised(ci, 1.0, 1.5), tp.fresnel_unpolarised(ct, 1.5, 1.0), atol=1e-12)


def test_furnace_series_of_lambert_is_geometric():
    # With a pure Lambert BSDF the 1-D transport gives the textbook furnace
    # sum a^k; the quadrature of T reproduces it.
    sc = tp.FurnaceSphere.__new__(tp.FurnaceSphere)
    real = tp.opaque_bsdf
    try:
        tp.opaque_bsdf = lambda wi, wo, n, *a: (np.broadcast_to(
            np.where((np.sum(wi * n, -1) > 0)[..., None], 0.6 / np.pi, 0.0), np.broadcast_shapes(wi.shape, wo.shape)),
            None)
        got = sc.series(np.array([0.3, 0.9]), 4)
    finally:
        tp.opaque_bsdf = real
    assert np.allclose(got, sum(0.6 ** k for k in range(5)), rtol=1e-10)


@pytest.mark.parametrize("candidates", [1, 8])
def test_direct_light_nee_and_ris(candidates):
    pc.direct_light_case(ORC, candidates=candidates)


def test_direct_light_restir_primary():
    pc.direct_light_case(ORC, candidates=8, restir=True)


def test_direct_light_mis_depth1():
    pc.direct_light_case(ORC, nomis=0)


def test_direct_light_mis_depth2():
    pc.direct_light_case(ORC, depth=2, nomis=0)


def test_direct_light_f64():
    pc.direct_light_case(ORC64, spp=256)
    pc.direct_light_case(ORC64, depth=2, nomis=0, spp=256)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_furnace(depth):
    pc.furnace_case(ORC, depth)


def test_furnace_restir_and_f64():
    pc.furnace_case(ORC, 3, restir=True, candidates=4)
    pc.furnace_case(ORC64, 3, spp=256)


@pytest.mark.parametrize("depth", [2, 3])
def test_glass_slab(depth):
    pc.slab_case(ORC, depth)


def test_glass_slab_restir_and_f64():
    pc.slab_case(ORC, 3, restir=True)
    pc.slab_case(ORC64, 3, spp=256)


def test_hue_shader():
    pc.hue_case(ORC)


@pytest.mark.parametrize("case", pc.ALL_FILTER_CASES, ids=lambda c: c.__name__)
def test_denoiser_closed_forms(case):
    case(ORC)


@pytest.mark.parametrize("case", pc.ALL_MERGE_CASES, ids=lambda c: c.__name__)
def test_restir_merge_closed_forms(case):
    case(ORC)
