#!/usr/bin/env python3
"""Benchmark of the HIP rendering backend on one MI355X.

Workload (one step = one frame): the mixed-material sphere scene (64
icospheres of subdivision 4 + ground + 8 emissive quads, 327,698 triangles,
seed 0) at 1280x720, 16 spp, max depth 8, FP32, through the ReSTIR DI path
(32 RIS candidates, temporal reuse + 3 spatial neighbours within 30 px) and the
edge-aware denoiser with 5 a-trous passes; the camera orbits at radius 10 over
34 frames with a static BVH, frame f renders with seed 12345 + f. The scene is
resident on the GPU before the timed region; each frame is one synchronous
rm_render call including the 2.7 MB RGB8 readback.

Prints one JSON line: Mrays/s (rays actually traced, counted on the GPU),
ms/frame, the roofline of the dominant kernel measured with HIP events, and
the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import raymotion_amd as rm  # noqa: E402
from raymotion_amd import scenes  # noqa: E402

METRIC = "Mrays/s & ms/frame 720p 16spp; ms to equal RMSE CNN vs SVGF; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0
PASS_NAMES = ["megakernel", "restir_generate", "restir_spatial", "restir_shade", "denoise_temporal",
              "denoise_atrous", "tonemap"]
# Algorithmic bytes: a 4-wide node is one 128-B line, a triangle test reads
# 48 B of leaf-ordered vertices; per-pixel buffers are counted once per pass.
NODE_BYTES = 128
TRI_BYTES = 48
PIXEL_BYTES = {  # per pixel and launch: reads + writes of the pass's own buffers
    "restir_generate": 80 + 64,            # G-buffer write + initial reservoir write
    "restir_spatial": 80 + 64 + 64,        # G-buffer + initial reservoir read, merged write
    "restir_shade": 80 + 64 + 16,          # G-buffer + merged reservoir read, radiance write
    "megakernel": 3 + 16,
}


def frame_settings(frame_index: int) -> rm.Frame:
    f = rm.Frame()
    f.width, f.height, f.spp, f.bounces = 1280, 720, 16, 8
    f.resolution, f.clamp_value = 1.0, 16.0
    f.adapt_min, f.adapt_step, f.adapt_rel, f.adapt_abs = 16, 8, 0.02, 0.002
    f.nomis = 0
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = 32, 3, 128, 30.0
    f.denoise, f.denoise_iterations, f.temporal = 1, 5, 1
    f.camera = scenes.mixed_camera(frame_index % 34)
    return f


def cpu_baseline(exported) -> dict:
    """The CPU oracle on a bounded sample: the full 1280x720 frame 0 of the same
    scene and settings at 1 spp (instead of 16), all host threads up to 16."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bindings as ob
    threads = min(16, os.cpu_count() or 1)
    f = frame_settings(0)
    f.spp = 1
    t0 = time.perf_counter()
    _, _, info = ob.render_frame(exported, f, 12345, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(info["rays"] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle, full 1280x720 frame 0 at 1 spp (16 on the GPU), same scene and ReSTIR/denoise "
                      f"settings: {info['rays']} rays in {dt:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel (scales HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        return 2
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 3}))
        return 0

    scene = scenes.mixed_spheres(seed=0)
    scene.build_bvh()
    exported = scene.export()
    backend = rm.Backend(0)
    backend.set_scene(exported)
    backend.set_profiling(True)

    for k in range(args.warmup):
        backend.render(frame_settings(k), 12345 + k, want_hdr=False)
    wall, rays, pass_ms, nodes, tris, prays = [], [], [], [], [], []
    for k in range(args.warmup, args.warmup + args.steps):
        f = frame_settings(k)
        t0 = time.perf_counter()
        backend.render(f, 12345 + k, want_hdr=False)
        wall.append((time.perf_counter() - t0) * 1e3)
        st = backend.stats()
        rays.append(st.rays[5])
        pass_ms.append([st.ms[i] for i in range(8)])
        nodes.append([st.node_visits[i] for i in range(6)])
        tris.append([st.tri_tests[i] for i in range(6)])
        prays.append([st.rays[i] for i in range(6)])
        if st.path != 1:
            raise RuntimeError("expected the ReSTIR + denoise path")
    backend.close()

    ms_per_frame = float(np.mean(wall))
    rays_per_frame = float(np.mean(rays))
    mrays = rays_per_frame / (ms_per_frame / 1e3) / 1e6
    pm = np.mean(np.array(pass_ms), axis=0)
    nd = np.mean(np.array(nodes, np.float64), axis=0)
    tr = np.mean(np.array(tris, np.float64), axis=0)
    ry = np.mean(np.array(prays, np.float64), axis=0)
    # Dominant kernel among the tracing passes (index = rm_stats pass index).
    tracing = {1: "restir_generate", 2: "restir_spatial", 3: "restir_shade"}
    dom = max(tracing, key=lambda i: pm[i])
    name = tracing[dom]
    px = 1280 * 720
    alg_bytes = nd[dom] * NODE_BYTES + tr[dom] * TRI_BYTES + px * PIXEL_BYTES[name]
    achieved = alg_bytes / (pm[dom] / 1e3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seed 0)",
        "config": {"workload": "mixed-material spheres, ReSTIR DI + a-trous denoiser, orbit camera",
                   "width": 1280, "height": 720, "spp": 16, "max_depth": 8, "triangles": int(len(exported.tris)),
                   "restir": {"candidates": 32, "spatial": 3, "radius_px": 30, "temporal": True},
                   "atrous_iterations": 5, "rays_per_frame": int(rays_per_frame),
                   "pass_ms": {PASS_NAMES[i]: round(float(pm[i]), 3) for i in range(1, 7)},
                   "pass_rays": {tracing[i]: int(ry[i]) for i in tracing},
                   "pass_nodes_per_ray": {tracing[i]: round(float(nd[i] / max(ry[i], 1.0)), 2) for i in tracing},
                   "pass_tris_per_ray": {tracing[i]: round(float(tr[i] / max(ry[i], 1.0)), 2) for i in tracing},
                   "gpu_frame_ms": round(float(pm[7]), 3)},
        "roofline": {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(alg_bytes), "kernel_ms": round(float(pm[dom]), 3)},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(exported)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
