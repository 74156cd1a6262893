"""The reference's known-answer GPU tests (tests/test_metal.cpp,
tests/test_alpha.cpp) through the HIP backend, plus bit-identity with the
oracle on the same cases."""
import numpy as np
import pytest

import kat_cases
import oracle_bindings as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", kat_cases.ALL, ids=lambda c: c.__name__)
def test_gpu_reference_kat(backend, case):
    def render(exported, f, seed):
        backend.set_scene(exported)
        rgb, _ = backend.render(f, seed, want_hdr=False)
        return rgb
    case(render)


def test_gpu_lit_scene_matches_oracle(backend):
    """The textured lit scene at the test's own settings (ReSTIR + denoiser,
    512 spp) in linear radiance at the 1e-3 bar."""
    from helpers import rel_err
    e, f = kat_cases.lit_textured_scene()
    backend.set_scene(e)
    _, g_hdr = backend.render(f, 12345)
    _, c_hdr, _ = ob.render_frame(e, f, 12345)
    err = rel_err(g_hdr, c_hdr)
    print(f"lit scene: max rel {err.max():.3g}, exact {np.mean(g_hdr == c_hdr):.6f}")
    assert err.max() <= 1e-3
