#!/usr/bin/env python3
"""Benchmark of the HIP rendering backend on one MI355X.

Workload (one step = one frame): the mixed-material sphere scene (64
icospheres of subdivision 4 + ground + 8 emissive quads, 327,698 triangles,
seed 0) at 1280x720, 16 spp, max depth 8, FP32, through the ReSTIR DI path
(32 RIS candidates, temporal reuse + 3 spatial neighbours within 30 px) and the
edge-aware denoiser with 5 a-trous passes; the camera orbits at radius 10 over
34 frames with a static BVH, frame f renders with seed 12345 + f. The scene is
resident on the GPU before the timed region; each frame is one synchronous
rm_render call including the 2.7 MB RGB8 readback. The same frames are then
timed through the drop-in boundary as well (render_image_hip's per-call
scene marshal and keying, then the render): boundary_ms_per_frame.

Prints one JSON line: Mrays/s (rays actually traced, counted on the GPU),
ms/frame, the roofline of the dominant kernel measured with HIP events, and
the CPU oracle timed on a bounded sample of the same workload.

With the default workload the line also carries "adjoint": the path-replay
adjoint's GPU time over the forward's on the 16-material inverse-rendering
scene at 512x512, 16 spp. --workload inverse runs the 200-step inverse
rendering (ms per iteration, loss and parameter error).

--workload heightfield measures the animated large-mesh case instead: a
1024x1024 height field (2,093,058 triangles + 256 emissive ones) displaced by
0.05 sin(2 pi t + x) every frame, 1920x1080, 8 spp, depth 6, ReSTIR defaults
(8 candidates, 4 spatial neighbours, radius 16) and the denoiser. Every frame
goes through the drop-in boundary: the host marshals the whole scene, the
backend uploads what changed and refits the resident tree on the GPU (a
rebuild from the host tree every 32nd frame), then renders. ms/frame is that boundary path;
the scene program's own vertex update and FP64 host refit are reported beside.

--workload rmse states each rendering config's image RMSE (linear and
tonemapped) against a 4096-spp path-traced reference of the same view, whose
64 launches of 64 spp are accumulated in FP64.
"""
from __future__ import annotations

import argparse
import json
import math
import re
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


# HBM bytes per launch come from rocprofv3 PMC runs of this same command
# (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's readings), committed as
# a table in this summary; bench.py cannot run the profiler itself.
PROFILE = ROOT / "profiles" / "restir_shade_c1_720p16spp.md"
# kernel names per pass in that table, any traversal-stack instance
PROFILE_KERNEL = {"restir_shade": r"k_persistent<\d+, PathEngine<\d+, true> >",
                  "restir_generate": r"k_persistent<\d+, GenEngine<\d+> >", "restir_spatial": r"k_spatial<\d+>"}


def measured_traffic(name: str):
    """HBM bytes (read + write) per launch of pass `name` from the committed PMC
    table, or None."""
    if not PROFILE.exists() or name not in PROFILE_KERNEL:
        return None
    rows = PROFILE.read_text().splitlines()
    start = next((i for i, l in enumerate(rows) if "FETCH KiB" in l), None)
    if start is None:
        return None
    for line in rows[start + 2:]:
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) != 5:
            break
        if re.fullmatch(PROFILE_KERNEL[name], cells[0]):
            return int((float(cells[3]) + float(cells[4])) * 1e6)
    return None


def roofline_entry(name: str, alg_bytes: float, kernel_ms: float) -> dict:
    """Roofline of the dominant tracing pass: algorithmic bytes (node and
    triangle fetches counted on the GPU, pixel buffers) per launch over its
    HIP-event time, against the HBM peak. The HBM bytes the same launch
    actually moved come from rocprofv3 PMC runs of this command (committed
    summary): they are a fraction of the algorithmic bytes because the tree and
    vertices stay in the 256 MB Infinity Cache, so the pass is limited by the
    latency of its dependent node fetches, which the measured HBM rate shows."""
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    traffic = measured_traffic(name)
    entry = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": int(alg_bytes),
             "kernel_ms": round(kernel_ms, 3), "limiter": "latency (dependent node fetches; tree cache-resident)"}
    if traffic:
        hbm = traffic / (kernel_ms / 1e3) / 1e9
        entry["measured_hbm_gbs"] = round(hbm, 1)
        entry["measured_hbm_frac"] = round(hbm / HBM_PEAK_GBS, 4)
        entry["traffic_source"] = f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this command, {PROFILE.name}"
    return entry


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


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """Threads for the CPU baseline: the cores this process may run on, up to
    16 (the host share of one GPU on the measurement boxes)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(exported, f: rm.Frame, what: str) -> dict:
    """The CPU oracle on a bounded sample: one full frame of the same scene and
    settings at 1 spp instead of the GPU's count, on the host's threads and on
    one thread."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bindings as ob
    threads = host_threads()
    gpu_spp = f.spp
    f.spp = 1
    t0 = time.perf_counter()
    _, _, info = ob.render_frame(exported, f, 12345, threads=threads)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, info1 = ob.render_frame(exported, f, 12345, threads=1)
    dt1 = time.perf_counter() - t0
    return {"value": round(info["rays"] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "single_thread": round(info1["rays"] / dt1 / 1e6, 3), "cpu": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle, full {f.width}x{f.height} {what} at 1 spp ({gpu_spp} on the GPU), same scene and "
                      f"ReSTIR/denoise settings: {info['rays']} rays in {dt:.2f} s on {threads} threads, "
                      f"{dt1:.2f} s on one"}


def heightfield_frame(frame_index: int) -> rm.Frame:
    f = rm.Frame()
    f.width, f.height, f.spp, f.bounces = 1920, 1080, 8, 6
    f.resolution, f.clamp_value = 1.0, 16.0
    f.adapt_min, f.adapt_step, f.adapt_rel, f.adapt_abs = 16, 8, 0.02, 0.002
    f.nomis = 0
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = 8, 4, 128, 16.0
    f.denoise, f.denoise_iterations, f.temporal = 1, 2, 1
    f.camera = scenes.heightfield_camera(frame_index)
    return f


def run_heightfield(args) -> dict:
    """Animated height field through the boundary; frame f is at t = f / 60 and
    the tree is rebuilt when f % 32 == 0 (Objects::render's policy), on the GPU
    (rm_set_scene_gpu_tree; refit on the GPU otherwise) or, with --host-tree,
    on the host (SAH build or FP64 refit, then the GPU refits its 4-wide copy)."""
    gpu_tree = not args.host_tree
    scene = scenes.heightfield(1024, seed=0)
    if not gpu_tree:
        scene.build_bvh()
    anim = scenes.HeightFieldAnimation(1024, seed=0)
    backend = rm.Backend(0)
    backend.set_scene(scene, gpu_tree=gpu_tree, rebuild=True)
    backend.set_profiling(True)

    def one(f_index):
        pos = anim.positions(f_index / 60.0)
        t0 = time.perf_counter()
        scene.set_positions(pos)
        if not gpu_tree:
            if f_index % 32 == 0:
                scene.build_bvh()
            else:
                scene.refit_bvh()
        t1 = time.perf_counter()
        # rm_scene_upload(_gpu_tree): marshal to 176-B records, then the upload entry point
        backend.set_scene(scene, gpu_tree=gpu_tree, rebuild=f_index % 32 == 0)
        t3 = time.perf_counter()
        backend.render(heightfield_frame(f_index), 12345 + f_index, want_hdr=False)
        t4 = time.perf_counter()
        st = backend.stats()
        return {"update": t1 - t0, "upload": t3 - t1, "render": t4 - t3,
                "boundary": t4 - t1, "rays": st.rays[5], "refit": st.refit, "gpu_ms": st.ms[7],
                "upload_ms": st.scene_ms[0], "refit_ms": st.scene_ms[1], "pass_ms": [st.ms[i] for i in range(8)],
                "nodes": [st.node_visits[i] for i in range(6)], "tris": [st.tri_tests[i] for i in range(6)],
                "prays": [st.rays[i] for i in range(6)], "bvh_nodes": st.bvh_nodes, "stack": st.stack_depth}

    for k in range(1, 1 + args.warmup):
        one(k)
    recs = [one(k) for k in range(1 + args.warmup, 1 + args.warmup + args.steps)]
    rebuild = one(32)
    backend.close()
    if any(r["refit"] != 1 for r in recs):
        raise RuntimeError(f"expected GPU refits, got {[r['refit'] for r in recs]}")
    mean = {k: float(np.mean([r[k] for r in recs])) for k in ("update", "upload", "render", "boundary",
                                                               "rays", "gpu_ms", "upload_ms", "refit_ms")}
    pm = np.mean(np.array([r["pass_ms"] for r in recs]), axis=0)
    nd = np.mean(np.array([r["nodes"] for r in recs], np.float64), axis=0)
    tr = np.mean(np.array([r["tris"] for r in recs], np.float64), axis=0)
    ry = np.mean(np.array([r["prays"] for r in recs], np.float64), axis=0)
    tracing = {1: "restir_generate", 2: "restir_spatial", 3: "restir_shade"}
    dom = max(tracing, key=lambda i: pm[i])
    name = tracing[dom]
    px = 1920 * 1080
    alg_bytes = nd[dom] * NODE_BYTES + tr[dom] * TRI_BYTES + px * PIXEL_BYTES[name]
    achieved = alg_bytes / (pm[dom] / 1e3) / 1e9
    ntri = scene.counts()[0]
    nnodes = recs[0]["bvh_nodes"]
    # refit: every triangle's 48 B read once, every node written and read once by its parent
    refit_bytes = ntri * TRI_BYTES + 2 * nnodes * NODE_BYTES
    ms_per_frame = mean["boundary"] * 1e3
    # over a 32-frame cycle, the host scene update (and host tree work) included
    amortized = (31 * (mean["boundary"] + mean["update"]) + rebuild["boundary"] + rebuild["update"]) * 1e3 / 32
    line = {
        "metric": METRIC,
        "value": round(mean["rays"] / mean["boundary"] / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural height field, seed 0)",
        "config": {"workload": "animated 1024x1024 height field, 256 emissive triangles, refit every frame, "
                               "rebuild every 32", "width": 1920, "height": 1080, "spp": 8, "max_depth": 6,
                   "triangles": ntri, "restir": {"candidates": 8, "spatial": 4, "radius_px": 16, "temporal": True},
                   "atrous_iterations": 2, "rays_per_frame": int(mean["rays"]),
                   "tree": "built on the GPU (rebuilt every 32nd frame, GPU refit otherwise)" if gpu_tree else
                           "host SAH (rebuilt every 32nd frame, FP64 host refit otherwise; GPU refits its copy)",
                   "host_ms": {"scene_update" + ("" if gpu_tree else "_and_fp64_refit"): round(mean["update"] * 1e3, 1),
                               "marshal_and_set_scene": round(mean["upload"] * 1e3, 1),
                               "render_call": round(mean["render"] * 1e3, 1)},
                   "gpu_ms": {"upload_copies": round(mean["upload_ms"], 3), "refit_kernels": round(mean["refit_ms"], 3),
                              "frame_passes": round(mean["gpu_ms"], 3)},
                   "gpu_mrays_per_s": round(mean["rays"] / (mean["gpu_ms"] / 1e3) / 1e6, 2),
                   "refit_gbs": round(refit_bytes / (mean["refit_ms"] / 1e3) / 1e9, 1) if mean["refit_ms"] > 0 else None,
                   "tree_nodes": rebuild["bvh_nodes"], "tree_stack_depth": rebuild["stack"],
                   "rebuild_frame_ms": round((rebuild["boundary"] + rebuild["update"]) * 1e3, 1),
                   "rebuild_frame_tree": {0: "unchanged", 1: "refit (same topology)", 2: "rebuilt from the host tree",
                                          3: "built on the GPU"}[rebuild["refit"]],
                   "rebuild_tree_ms": round((rebuild["update"] if not gpu_tree else 0.0) * 1e3 + rebuild["refit_ms"], 1),
                   "amortized_ms_per_frame_with_scene_update": round(amortized, 1),
                   "pass_ms": {PASS_NAMES[i]: round(float(pm[i]), 3) for i in range(1, 7)},
                   "pass_rays": {tracing[i]: int(ry[i]) for i in tracing},
                   "pass_nodes_per_ray": {tracing[i]: round(float(nd[i] / max(ry[i], 1.0)), 2) for i in tracing}},
        "roofline": {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(alg_bytes), "kernel_ms": round(float(pm[dom]), 3)},
    }
    if not args.no_cpu_baseline:
        if gpu_tree:
            scene.build_bvh()  # the oracle traverses the host tree
        line["cpu_baseline"] = cpu_baseline(scene.export(), heightfield_frame(32), "heightfield frame 32")
    return line


def cornell_frame_c0() -> rm.Frame:
    """The Cornell box frame: 256x256, 4 spp, depth 5, NEE+MIS, ReSTIR and
    denoiser off; pinhole at (0.5, 0.5, -1.4) with a 40 degree vertical field
    of view (focal = (H/2) / tan(20 deg))."""
    f = rm.Frame()
    f.width, f.height, f.spp, f.bounces = 256, 256, 4, 5
    f.resolution, f.clamp_value = 1.0, 16.0
    f.adapt_min, f.adapt_step, f.adapt_rel, f.adapt_abs = 16, 8, 0.02, 0.002
    f.nomis = 0
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = 1, 0, 128, 16.0
    f.denoise, f.denoise_iterations, f.temporal = 0, 0, 0
    cam = scenes.cornell_camera()
    cam.focal = 128.0 / math.tan(math.radians(20.0))
    f.camera = cam
    return f


def reference_of(f: rm.Frame) -> rm.Frame:
    """Path-traced reference settings for frame f: same size, camera, depth and
    clamp; legacy light sampling, no reuse, no denoiser, no adaptive stop."""
    r = rm.Frame.from_buffer_copy(f)
    r.spp = REF_BATCH_SPP
    r.adapt_min = 0
    r.restir_candidates, r.restir_spatial = 1, 0
    r.denoise, r.temporal = 0, 0
    return r


REF_SPP, REF_BATCH_SPP = 4096, 64


def reference_image(backend, f: rm.Frame, seed0: int):
    """REF_SPP samples per pixel as REF_SPP / REF_BATCH_SPP launches with
    distinct seeds; the launches' per-pixel means are accumulated in FP64.
    Returns the mean and its standard error per pixel channel."""
    r = reference_of(f)
    n = REF_SPP // REF_BATCH_SPP
    acc = None
    for k in range(n):
        _, hdr = backend.render(r, seed0 + 7919 * k)
        x = hdr.astype(np.float64)
        if acc is None:
            acc, acc2 = np.zeros_like(x), np.zeros_like(x)
        acc += x
        acc2 += x * x
    mean = acc / n
    stderr = np.sqrt(np.maximum(acc2 / n - mean * mean, 0.0) / (n - 1))
    return mean, stderr


def image_errors(img, ref, ref_stderr) -> dict:
    from raymotion_amd.denoise import rmse, tonemapped_rmse
    return {"rmse": round(rmse(img, ref), 6), "rmse_tonemapped": round(tonemapped_rmse(img, ref), 6),
            "reference_noise_rms": round(float(np.sqrt(np.mean(ref_stderr ** 2))), 6)}


def run_rmse(args) -> dict:
    """Each rendering config's output frame against a 4096-spp path-traced
    reference of the same view accumulated in FP64: the Cornell box (4 spp
    megakernel), the mixed scene (frame 4 of the orbit after 4 frames of
    history, 16 spp ReSTIR + 5 a-trous passes), the animated height field
    (frame 4 after 4 frames, 8 spp ReSTIR + 2 a-trous passes, GPU tree), the
    inverse-rendering scene at its ground truth (512x512, 16 spp path tracing)
    and the learned denoiser (mixed-scene frame 4, 4 spp ReSTIR + the network
    with the committed weights)."""
    out = {}
    t0 = time.perf_counter()
    backend = rm.Backend(0)
    # Cornell box
    s = scenes.cornell_box()
    s.build_bvh()
    backend.set_scene(s)
    f = cornell_frame_c0()
    _, img = backend.render(f, 12345)
    ref, se = reference_image(backend, f, 500_000)
    out["cornell_256_4spp"] = image_errors(img, ref, se)
    # mixed scene, orbit frames 0..4
    s = scenes.mixed_spheres(seed=0)
    s.build_bvh()
    backend.set_scene(s, gpu_tree=True, rebuild=True)
    for k in range(5):
        f = frame_settings(k)
        f.temporal = 0 if k == 0 else 1
        _, img = backend.render(f, 12345 + k)
    ref, se = reference_image(backend, f, 600_000)
    out["mixed_720p_16spp_restir_atrous"] = image_errors(img, ref, se)
    # animated height field, frames 0..4
    s = scenes.heightfield(1024, seed=0)
    anim = scenes.HeightFieldAnimation(1024, seed=0)
    for k in range(5):
        s.set_positions(anim.positions(k / 60.0))
        backend.set_scene(s, gpu_tree=True, rebuild=k == 0)
        f = heightfield_frame(k)
        f.temporal = 0 if k == 0 else 1
        _, img = backend.render(f, 12345 + k)
    ref, se = reference_image(backend, f, 700_000)
    out["heightfield_1080p_8spp_restir_atrous"] = image_errors(img, ref, se)
    # inverse-rendering scene at its ground-truth parameters: the 16-spp
    # path-tracing frame the loss is taken on
    from raymotion_amd.diff import DiffScene, path_frame
    s, groups = scenes.inverse_scene(seed=0)
    ds = DiffScene(s.export(), groups)
    backend.set_scene(ds.exported(ds.initial))
    f = path_frame(512, 512, 16, 8, scenes.mixed_camera(0))
    _, img = backend.render(f, 12345)
    ref, se = reference_image(backend, f, 800_000)
    out["inverse_512_16spp_path"] = image_errors(img, ref, se)
    # learned denoiser at 4 spp (mixed-scene orbit frame 4 after 4 frames)
    if WEIGHTS.exists():
        from raymotion_amd import denoise
        s = scenes.mixed_spheres(seed=0)
        s.build_bvh()
        backend.set_scene(s, gpu_tree=True, rebuild=True)
        backend.kpn_set_weights(denoise.load_weights(WEIGHTS))
        for k in range(5):
            f = denoise.restir_frame(1280, 720, 4, scenes.mixed_camera(k), denoise=2, temporal=0 if k == 0 else 1)
            _, img = backend.render(f, 12345 + k)
        ref, se = reference_image(backend, f, 900_000)
        out["mixed_720p_4spp_restir_learned_denoiser"] = image_errors(img, ref, se)
    backend.close()
    return {"metric": "image RMSE against a 4096-spp FP64-accumulated path-traced reference", "value": None,
            "unit": "linear radiance RMSE", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": None,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (procedural scenes, seed 0)",
            "config": {"workload": "per-config image error", "reference_spp": REF_SPP,
                       "reference_batches": REF_SPP // REF_BATCH_SPP, "seconds": round(time.perf_counter() - t0, 1)},
            "errors": out}


def adjoint_cost(steps: int = 5) -> dict:
    """Forward vs path-replay adjoint GPU time on the 16-material inverse-rendering
    scene (the mixed generator, 327,698 triangles) at 512x512, 16 spp, depth 8."""
    from raymotion_amd.diff import DiffScene, Renderer, path_frame
    s, groups = scenes.inverse_scene(seed=0)
    ds = DiffScene(s.export(), groups)
    backend = rm.Backend(0)
    backend.set_profiling(True)
    r = Renderer(backend, ds, path_frame(512, 512, 16, 8, scenes.mixed_camera(0)))
    d = np.random.default_rng(0).uniform(0, 1, 512 * 512 * 3).astype(np.float32)
    fwd, adj, fwd_wall, adj_wall, rays_f, rays_a = [], [], [], [], [], []
    for k in range(steps + 1):
        t0 = time.perf_counter()
        r.forward(ds.initial, 7 + k)
        t1 = time.perf_counter()
        st = backend.stats()
        f_ms, f_rays = st.ms[7], st.rays[5]
        r.adjoint(ds.initial, 7 + k, d)
        t2 = time.perf_counter()
        st = backend.stats()
        if k:  # the first pair warms up
            fwd.append(f_ms); adj.append(st.ms[7]); rays_f.append(f_rays); rays_a.append(st.rays[5])
            fwd_wall.append((t1 - t0) * 1e3); adj_wall.append((t2 - t1) * 1e3)
    backend.close()
    return {"workload": "16-material mixed scene, 512x512, 16 spp, depth 8, path-tracing frame",
            "parameters": int(ds.ngroup * 5), "forward_ms": round(float(np.mean(fwd)), 3),
            "adjoint_ms": round(float(np.mean(adj)), 3), "ratio": round(float(np.mean(adj) / np.mean(fwd)), 3),
            "forward_rays": int(np.mean(rays_f)), "adjoint_rays": int(np.mean(rays_a)),
            "forward_wall_ms": round(float(np.mean(fwd_wall)), 2), "adjoint_wall_ms": round(float(np.mean(adj_wall)), 2)}


def run_inverse(args) -> dict:
    """Inverse rendering on the 16-material scene: target at 512x512, 64 spp from
    the ground-truth parameters; start from the ground truth perturbed by
    U[0.7, 1.3] (seed 2); Adam, lr 1e-2, pixel L2 at 16 spp."""
    from raymotion_amd.diff import DiffScene, Renderer, optimize, path_frame
    s, groups = scenes.inverse_scene(seed=0)
    ds = DiffScene(s.export(), groups)
    backend = rm.Backend(0)
    cam = scenes.mixed_camera(0)
    r = Renderer(backend, ds, path_frame(512, 512, 16, 8, cam))
    hi = Renderer(backend, ds, path_frame(512, 512, 64, 8, cam))
    gt = ds.initial
    target = hi.forward(gt, 1)
    p0 = ds.clamp(gt * np.random.default_rng(2).uniform(0.7, 1.3, gt.shape))
    steps = args.steps
    t0 = time.perf_counter()
    p, losses = optimize(r, target, p0, steps=steps, lr=1e-2)
    wall = time.perf_counter() - t0

    def blocks(img, k=8):
        return img.reshape(512 // k, k, 512 // k, k, 3).mean(axis=(1, 3))

    def image_error(q):
        return float(np.mean((blocks(hi.forward(q, 2)) - blocks(target)) ** 2))

    e0, e1, floor = image_error(p0), image_error(p), image_error(gt)
    g = np.abs(r.adjoint(gt, 3, np.ones(512 * 512 * 3, np.float32)))
    backend.close()
    strong = g >= 0.1 * g.max()
    observed = g > 0
    return {
        "metric": METRIC, "value": round(wall / steps * 1e3, 2), "unit": "ms/iteration", "n_gpus": 1,
        "steps": steps, "warmup": 0, "ms_per_step": round(wall / steps * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scene, seed 0)",
        "config": {"workload": "inverse rendering, 16-material mixed scene, 80 parameters", "width": 512,
                   "height": 512, "spp": 16, "target_spp": 64, "max_depth": 8, "optimizer": "Adam lr 1e-2",
                   "loss_first5": round(float(losses[:5].mean()), 6), "loss_last5": round(float(losses[-5:].mean()), 6),
                   "block8_image_error": {"start": e0, "end": e1, "ground_truth": floor},
                   "param_abs_error_observed": {"start": round(float(np.abs(p0 - gt)[observed].mean()), 4),
                                                "end": round(float(np.abs(p - gt)[observed].mean()), 4)},
                   "param_abs_error_strong": {"count": int(strong.sum()),
                                              "start": round(float(np.abs(p0 - gt)[strong].mean()), 4),
                                              "end": round(float(np.abs(p - gt)[strong].mean()), 4)}},
    }


WEIGHTS = ROOT / "raymotion_amd" / "data" / "kpn_mixed.npy"


def run_denoiser_train(args) -> dict:
    """Trains the learned denoiser: 512 random scenes of the mixed-scene
    generator, one 256x256 crop each (4-spp ReSTIR input, 4096-spp
    path-traced reference), 20,000 Adam steps at lr 1e-3 and batch 16, then
    evaluates it against the a-trous path."""
    from raymotion_amd import denoise
    backend = rm.Backend(0)
    backend.set_profiling(True)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    data_s = denoise.make_training_set(backend, args.views, size=256, noisy_spp=4, ref_spp=args.ref_spp,
                                       views_per_scene=1, seed=0, subdivisions=4, log=log)
    t0 = time.perf_counter()
    losses = denoise.train(backend, denoise.init_params(0), args.steps, batch=16, lr=1e-3, seed=1, log=log)
    train_s = time.perf_counter() - t0
    weights = backend.kpn_get_weights()
    if args.save_weights:
        np.save(args.save_weights, weights)
    ev = denoise.evaluate(backend, frames=(4, 20), svgf_spp=(4, 8, 16, 32, 64, 128, 256), cnn_spp=(1, 2, 4),
                          log=log)
    backend.close()
    return {"metric": METRIC, "value": round(train_s / max(args.steps, 1) * 1e3, 3), "unit": "ms/training step",
            "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": round(train_s / max(args.steps, 1) * 1e3, 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic (random mixed scenes, rendered by the backend)",
            "config": {"workload": "kernel-predicting denoiser training", "views": args.views, "crop": 256,
                       "scenes": args.views, "input_spp": 4, "reference_spp": args.ref_spp, "batch": 16,
                       "optimizer": "Adam lr 1e-3",
                       "training_set_s": round(data_s, 1), "training_s": round(train_s, 1),
                       "loss_first100": float(np.mean(losses[:100])), "loss_last100": float(np.mean(losses[-100:]))},
            "denoiser": ev}


def denoiser_extras() -> dict | None:
    """Equal-RMSE comparison with the committed trained weights (frames 4 and 20
    of the mixed-scene orbit, 4096-spp references)."""
    from raymotion_amd import denoise
    if not WEIGHTS.exists():
        return None
    backend = rm.Backend(0)
    backend.kpn_set_weights(denoise.load_weights(WEIGHTS))
    ev = denoise.evaluate(backend, frames=(4, 20), svgf_spp=(4, 8, 16, 32, 64, 128, 256), cnn_spp=(1, 2, 4))
    backend.close()
    return ev


# Multiply-adds of the network per pixel (13 real input channels, 25 logits),
# times 2 FLOP: 2 * 9 * (13*64 + 4*64*64 + 64*25).
KPN_FLOP_PER_PIXEL = 2 * 9 * (13 * 64 + 4 * 64 * 64 + 64 * 25)
BF16_PEAK_TFLOPS = 2500.0


def run_denoiser_infer(args) -> dict:
    """Frames of the mixed-scene orbit through ReSTIR at 4 spp and the learned
    denoiser (committed weights) at 1280x720; the network's MFMA roofline."""
    from raymotion_amd import denoise
    scene = scenes.mixed_spheres(seed=0)
    scene.build_bvh()
    backend = rm.Backend(0)
    backend.set_scene(scene)
    backend.set_profiling(True)
    backend.kpn_set_weights(denoise.load_weights(WEIGHTS))
    frame_ms, net_ms = [], []
    for k in range(args.warmup + args.steps):
        backend.render(denoise.restir_frame(1280, 720, 4, scenes.mixed_camera(k % 34), denoise=2), 12345 + k,
                       want_hdr=False)
        st = backend.stats()
        if k >= args.warmup:
            frame_ms.append(st.ms[7])
            net_ms.append(st.ms[5])
    backend.close()
    ms = float(np.mean(net_ms))
    flop = KPN_FLOP_PER_PIXEL * 1280 * 720
    tflops = flop / (ms / 1e3) / 1e12
    return {"metric": METRIC, "value": round(float(np.mean(frame_ms)), 3), "unit": "ms/frame", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(float(np.mean(frame_ms)), 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic (procedural scene, seed 0)",
            "config": {"workload": "mixed-scene orbit, ReSTIR 4 spp + learned KPN denoiser, 1280x720",
                       "network_ms": round(ms, 3)},
            "roofline": {"kernel": "kpn network (6 convolutions, the 5x5 filter in the head's epilogue)", "bound": "mfma", "achieved": round(tflops, 1),
                         "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / BF16_PEAK_TFLOPS, 4),
                         "traffic": None, "flop_per_launch": flop}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed frames (10; heightfield 8)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed frames first (2; heightfield 1)")
    ap.add_argument("--workload", choices=("mixed", "heightfield", "inverse", "denoiser-train", "denoiser-infer", "rmse"),
                    default="mixed")
    ap.add_argument("--views", type=int, default=512, help="denoiser-train: training views")
    ap.add_argument("--ref-spp", type=int, default=4096, help="denoiser-train: reference samples per pixel")
    ap.add_argument("--save-weights", default=None, help="denoiser-train: write the trained weights (.npy)")
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint cost measurement in the mixed line")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    ap.add_argument("--host-tree", action="store_true",
                    help="upload the engine's host SAH tree (and, heightfield, refit it on the host) instead of "
                         "building the traversal tree on the GPU")
    ap.add_argument("--schedule", type=int, choices=(0, 1, 2), default=0,
                    help="path passes: 0 persistent lanes, 1 wavefront + persistent tail, 2 wavefront only")
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

    hf = args.workload == "heightfield"
    if args.steps is None:
        args.steps = {"heightfield": 8, "inverse": 200, "denoiser-train": 20000}.get(args.workload, 10)
    if args.warmup is None:
        args.warmup = 1 if hf else 2
    if args.workload == "inverse":
        line = run_inverse(args)
    elif args.workload == "denoiser-train":
        line = run_denoiser_train(args)
    elif args.workload == "denoiser-infer":
        line = run_denoiser_infer(args)
    elif args.workload == "rmse":
        line = run_rmse(args)
    else:
        line = run_heightfield(args) if hf else run_mixed(args)
        if args.workload == "mixed" and not args.no_extras:
            line["adjoint"] = adjoint_cost()
            line["denoiser"] = denoiser_extras()
    print(json.dumps(line), flush=True)
    return 0


def run_mixed(args) -> dict:
    scene = scenes.mixed_spheres(seed=0)
    scene.build_bvh()
    exported = scene.export()
    backend = rm.Backend(0)
    backend.set_scene(exported, gpu_tree=not args.host_tree, rebuild=True)
    backend.set_profiling(True)
    backend.set_schedule(args.schedule)

    for k in range(args.warmup):
        backend.render(frame_settings(k), 12345 + k, want_hdr=False)
    wall, rays, pass_ms, nodes, tris, prays, waves = [], [], [], [], [], [], []
    for k in range(args.warmup, args.warmup + args.steps):
        f = frame_settings(k)
        t0 = time.perf_counter()
        backend.render(f, 12345 + k, want_hdr=False)
        wall.append((time.perf_counter() - t0) * 1e3)
        st = backend.stats()
        waves.append((st.wave_iters, st.wave_tail))
        tree_shape = (st.bvh_nodes, st.stack_depth)
        rays.append(st.rays[5])
        pass_ms.append([st.ms[i] for i in range(8)])
        nodes.append([st.node_visits[i] for i in range(6)])
        tris.append([st.tri_tests[i] for i in range(6)])
        prays.append([st.rays[i] for i in range(6)])
        if st.path != 1:
            raise RuntimeError("expected the ReSTIR + denoise path")
    # The same frames through the drop-in boundary: render_image_hip marshals
    # the Scene to 176-byte records and re-keys it every call (unchanged here,
    # so nothing is re-sent), then renders.
    boundary, resent = [], 0
    if not args.host_tree:
        backend.set_scene(scene, gpu_tree=True)  # untimed: the first marshal allocates its staging
        for k in range(args.warmup, args.warmup + args.steps):
            t0 = time.perf_counter()
            backend.set_scene(scene, gpu_tree=True)
            backend.render(frame_settings(k), 12345 + k, want_hdr=False)
            boundary.append((time.perf_counter() - t0) * 1e3)
            resent += backend.stats().refit != 0
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
                   "gpu_frame_ms": round(float(pm[7]), 3),
                   "boundary_ms_per_frame": round(float(np.mean(boundary)), 3) if boundary else None,
                   "boundary_mrays_per_s": round(rays_per_frame / (float(np.mean(boundary)) / 1e3) / 1e6, 2)
                   if boundary else None,
                   "boundary_frames_resending_the_scene": resent if boundary else None,
                   "tree": "host SAH (the scene's own)" if args.host_tree else "built on the GPU",
                   "tree_nodes": tree_shape[0], "tree_stack_depth": tree_shape[1],
                   "schedule": {0: "persistent", 1: "wavefront", 2: "wavefront-only"}[args.schedule],
                   "wavefront_iterations": round(float(np.mean([w[0] for w in waves])), 1),
                   "wavefront_tail_pixels": int(np.mean([w[1] for w in waves]))},
        "roofline": roofline_entry(name, alg_bytes, float(pm[dom])),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(exported, frame_settings(0), "frame 0")
    return line


if __name__ == "__main__":
    sys.exit(main())
