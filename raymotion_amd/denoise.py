"""Learned kernel-predicting denoiser (KPN) for the ReSTIR frame.

The network is six 3x3 convolutions (16 -> 64, four 64 -> 64 with ReLU,
64 -> 25 logits) over per-pixel features of the ReSTIR frame; a softmax of
the logits is a 5x5 filter applied to the noisy radiance. Forward, backward
and Adam run on the GPU in BF16 MFMA kernels with FP32 accumulation
(librm_hip); this module holds the parameter layout, initialisation, the
training-set generator (noisy ReSTIR frames paired with converged path-traced
references of random mixed-material scenes), training and evaluation.
The a-trous filter (frame.denoise = 1) stays the fallback.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import scenes
from ._capi import Frame, lib

FEATURES = 16
TAPS = 25
# (cin, cout) per layer, channel counts padded to multiples of 16
LAYERS = [(16, 64), (64, 64), (64, 64), (64, 64), (64, 64), (64, 32)]


@dataclass(frozen=True)
class Layer:
    cin: int
    cout: int
    kp: int      # 9 * cin rounded up to 32
    w_off: int   # [cout][kp] weights, K = tap * cin + ci, tap = 3 * ky + kx
    b_off: int   # cout biases


def layout() -> list[Layer]:
    out, off = [], 0
    for cin, cout in LAYERS:
        kp = (9 * cin + 31) // 32 * 32
        out.append(Layer(cin, cout, kp, off, off + cout * kp))
        off += cout * kp + cout
    return out


def param_count() -> int:
    n = sum(L.cout * L.kp + L.cout for L in layout())
    assert n == lib().rm_kpn_param_count(), "parameter layout differs from the library's"
    return n


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights on the real entries (zero padding), zero biases; the
    head starts near zero, i.e. near a 5x5 box filter."""
    rng = np.random.default_rng(seed)
    p = np.zeros(sum(L.cout * L.kp + L.cout for L in layout()), np.float32)
    for i, L in enumerate(layout()):
        w = np.zeros((L.cout, L.kp), np.float32)
        std = math.sqrt(2.0 / (9 * L.cin)) if i < len(LAYERS) - 1 else 1e-3
        cout = L.cout if i < len(LAYERS) - 1 else TAPS
        w[:cout, :9 * L.cin] = rng.normal(0.0, std, (cout, 9 * L.cin))
        p[L.w_off:L.w_off + L.cout * L.kp] = w.reshape(-1)
    return p


def restir_frame(width: int, height: int, spp: int, camera, denoise: int = 1, iterations: int = 5,
                 temporal: int = 1) -> Frame:
    """The mixed-scene ReSTIR settings: 32 candidates, 3 spatial neighbours in a
    30 px radius, depth 8."""
    f = Frame()
    f.width, f.height, f.spp, f.bounces = width, height, spp, 8
    f.resolution, f.clamp_value = 1.0, 16.0
    f.adapt_min, f.adapt_step, f.adapt_rel, f.adapt_abs = 16, 8, 0.02, 0.002
    f.nomis = 0
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = 32, 3, 128, 30.0
    f.denoise, f.denoise_iterations, f.temporal = denoise, iterations, temporal
    f.camera = camera
    return f


def reference_frame(width: int, height: int, spp: int, camera) -> Frame:
    """Path-traced reference: legacy light sampling, no reuse, no adaptive stop."""
    f = restir_frame(width, height, spp, camera, denoise=0)
    f.restir_candidates, f.restir_spatial = 1, 0
    f.adapt_min = 0
    return f


def random_view(rng) -> object:
    """A camera orbiting the mixed scene: radius 7-12, height 0.5-5, any angle."""
    return scenes.orbit_camera((0.0, 0.0, 0.0), rng.uniform(7.0, 12.0), rng.uniform(0.5, 5.0),
                               rng.uniform(0.0, 360.0), 45.0)


def make_training_set(backend, n_views: int, size: int = 256, noisy_spp: int = 4, ref_spp: int = 1024,
                      views_per_scene: int = 1, seed: int = 0, subdivisions: int = 4, log=None):
    """Renders n_views (features, noisy, reference) triples of random mixed
    scenes (a new scene every views_per_scene views; tree built on the GPU)
    into the backend's training set; returns seconds spent. A size x size
    frame of random_view's camera is the centre crop of the same camera's
    1280x720 frame: the focal length is fixed in pixels."""
    rng = np.random.default_rng(seed)
    backend.kpn_dataset_clear()
    t0 = time.perf_counter()
    for v in range(n_views):
        if v % views_per_scene == 0:
            s = scenes.mixed_spheres(seed=int(rng.integers(1 << 30)), subdivisions=subdivisions)
            backend.set_scene(s, gpu_tree=True, rebuild=True)
        cam = random_view(rng)
        feat, noisy = backend.render_features(restir_frame(size, size, noisy_spp, cam, denoise=0, temporal=0),
                                              int(rng.integers(1 << 62)))
        _, ref = backend.render(reference_frame(size, size, ref_spp, cam), int(rng.integers(1 << 62)))
        backend.kpn_dataset_add(feat, noisy, ref.reshape(size, size, 3))
        if log and (v + 1) % max(1, n_views // 8) == 0:
            log(f"training set: {v + 1}/{n_views} views, {time.perf_counter() - t0:.1f} s")
    return time.perf_counter() - t0


def train(backend, params0, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0, chunk: int = 500,
          log=None) -> np.ndarray:
    """Adam on the training set from params0; returns the per-step losses."""
    backend.kpn_set_weights(params0)
    losses = []
    t0 = time.perf_counter()
    for start in range(0, steps, chunk):
        n = min(chunk, steps - start)
        losses.append(backend.kpn_train(n, batch, lr, seed + start))
        if log:
            log(f"step {start + n}/{steps}: loss {float(np.mean(losses[-1][-50:])):.5f}, "
                f"{time.perf_counter() - t0:.1f} s")
    return np.concatenate(losses) if losses else np.zeros(0, np.float32)


def rmse(a, b) -> float:
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def tonemapped_rmse(a, b) -> float:
    a, b = np.maximum(np.asarray(a, np.float64), 0), np.maximum(np.asarray(b, np.float64), 0)
    return rmse(a / (a + 1), b / (b + 1))


def evaluate(backend, frames=(4,), ref_spp: int = 4096, history: int = 4, svgf_spp=(4, 8, 16, 32, 64),
             cnn_spp=(1, 2, 4), width: int = 1280, height: int = 720, log=None) -> dict:
    """RMSE against a ref_spp path-traced render and GPU ms/frame of the a-trous
    path (5 iterations) and the learned path on the mixed scene's orbit. Each
    evaluated frame f is the last of the sequence f - history .. f, so both
    paths have temporal history; ms/frame is the mean over the sequence.

    equal_rmse gives, per learned-path sample count, the a-trous cost at the
    same error (log-log interpolation over its sample counts) and the
    speed-up; where the a-trous sweep never reaches that error the cost of its
    largest sample count is a lower bound (marked)."""
    s = scenes.mixed_spheres(seed=0)
    s.build_bvh()
    backend.set_scene(s)
    backend.set_profiling(True)
    refs = {}
    for fr in frames:
        _, ref = backend.render(reference_frame(width, height, ref_spp, scenes.mixed_camera(fr)), 900_000 + fr)
        refs[fr] = ref
        if log:
            log(f"reference frame {fr}: {backend.stats().ms[7]:.0f} ms")

    def run(spp, mode):
        errs, terrs, ms = [], [], []
        for fr in frames:
            for k in range(fr - history, fr + 1):
                # each sequence starts without history (no carry-over from the previous run)
                f = restir_frame(width, height, spp, scenes.mixed_camera(k), denoise=mode,
                                 temporal=0 if k == fr - history else 1)
                _, hdr = backend.render(f, 12345 + k)
                ms.append(backend.stats().ms[7])
            errs.append(rmse(hdr, refs[fr]))
            terrs.append(tonemapped_rmse(hdr, refs[fr]))
        return {"spp": spp, "ms": round(float(np.mean(ms)), 3), "rmse": float(np.mean(errs)),
                "rmse_tonemapped": float(np.mean(terrs))}

    svgf = [run(spp, 1) for spp in svgf_spp]
    cnn = [run(spp, 2) for spp in cnn_spp]
    out = {"frames": list(frames), "reference_spp": ref_spp, "svgf": svgf, "cnn": cnn, "equal_rmse": {}}
    t = np.array([r["ms"] for r in svgf])
    for key in ("rmse", "rmse_tonemapped"):
        e = np.array([r[key] for r in svgf])
        rows = []
        for c in cnn:
            ms_eq = None
            if c[key] >= e[0]:
                ms_eq = float(t[0])  # matched at the lowest a-trous sample count swept
            for i in range(len(svgf) - 1):
                if ms_eq is None and e[i] >= c[key] >= e[i + 1]:
                    a = (math.log(c[key]) - math.log(e[i])) / (math.log(e[i + 1]) - math.log(e[i]))
                    ms_eq = float(math.exp(math.log(t[i]) + a * (math.log(t[i + 1]) - math.log(t[i]))))
            ms = ms_eq if ms_eq is not None else float(t[-1])
            rows.append({"cnn_spp": c["spp"], "svgf_ms": round(ms, 3), "cnn_ms": c["ms"],
                         "speedup": round(ms / c["ms"], 3), "lower_bound": ms_eq is None})
        out["equal_rmse"][key] = rows
    return out


def load_weights(path) -> np.ndarray:
    p = np.load(path, allow_pickle=False).astype(np.float32)
    if p.size != param_count():
        raise ValueError(f"{path}: {p.size} parameters, the network has {param_count()}")
    return p
