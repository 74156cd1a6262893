"""The transport, denoiser and ReSTIR-merge pins of test_pins_oracle.py run on
the HIP backend: each GPU result is checked against the expectation the
mathematics fixes (tests/transport_pins.py), with 8x the oracle's samples."""
import numpy as np
import pytest

import pin_cases as pc

pytestmark = pytest.mark.gpu


class GpuRunner:
    scale = 8
    batches = 32

    def __init__(self, backend):
        self.b = backend

    def hdr_batches(self, exported, frame, seeds):
        self.b.set_scene(exported)
        out = []
        for s in seeds:
            _, hdr = self.b.render(frame, s)
            out.append(hdr.reshape(-1, 3).copy())
        return np.stack(out)

    def gbuffer(self, exported, frame, seed):
        self.b.set_scene(exported)
        self.b.render(frame, seed)
        assert self.b.stats().path == 1
        g, _, _, _ = self.b.read_restir(frame.width, frame.height)
        return g

    def merge(self, exported, frame, seed, gbuf, rinit, hist_gbuf, hist_res, prev_camera):
        self.b.set_scene(exported)
        return self.b.restir_merge(frame, seed, gbuf, rinit, hist_gbuf, hist_res, prev_camera)

    def denoise(self, frame, colour, gbuf, hist_colour, hist_gbuf, prev_camera):
        return self.b.denoise_image(frame, colour, gbuf, hist_colour, hist_gbuf, prev_camera)


@pytest.fixture(scope="module")
def gpu(backend):
    return GpuRunner(backend)


@pytest.mark.parametrize("candidates", [1, 8])
def test_gpu_direct_light_nee_and_ris(gpu, candidates):
    print(pc.direct_light_case(gpu, candidates=candidates), flush=True)


def test_gpu_direct_light_restir_primary(gpu):
    print(pc.direct_light_case(gpu, candidates=8, restir=True), flush=True)


def test_gpu_direct_light_mis(gpu):
    print(pc.direct_light_case(gpu, nomis=0), flush=True)
    print(pc.direct_light_case(gpu, depth=2, nomis=0), flush=True)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_gpu_furnace(gpu, depth):
    print(pc.furnace_case(gpu, depth), flush=True)


def test_gpu_furnace_restir(gpu):
    print(pc.furnace_case(gpu, 3, restir=True, candidates=4), flush=True)


@pytest.mark.parametrize("depth", [2, 3])
def test_gpu_glass_slab(gpu, depth):
    print(pc.slab_case(gpu, depth), flush=True)
    print(pc.slab_case(gpu, depth, restir=True), flush=True)


def test_gpu_hue_shader(gpu):
    pc.hue_case(gpu)


@pytest.mark.parametrize("case", pc.ALL_FILTER_CASES, ids=lambda c: c.__name__)
def test_gpu_denoiser_closed_forms(gpu, case):
    case(gpu)


@pytest.mark.parametrize("case", pc.ALL_MERGE_CASES, ids=lambda c: c.__name__)
def test_gpu_restir_merge_closed_forms(gpu, case):
    case(gpu)
