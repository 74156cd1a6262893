"""Host scene engine and GPU backend, Python side.

Scene wraps the C++ engine (triangle precompute, SAH BVH, refit, marshalling
into the reference upload layout); Backend wraps one rm_ctx (device buffers,
render, traversal probe, intermediate readback). Names and argument meanings
follow the reference engine (src/engine/scene.cpp, bvh.cpp, metal.mm).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (GBUFFER_DTYPE, NODE_DTYPE, RESERVOIR_DTYPE, TRIANGLE_DTYPE, Camera, Frame, SceneDesc, Stats,
                    lib, ptr)


class BackendError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"rm error {code}: {message}")
        self.code = code


@dataclass
class Exported:
    """A scene in the reference upload layout (metal.mm:109-145)."""
    tris: np.ndarray
    nodes: np.ndarray
    bvh_tri: np.ndarray
    light_tri: np.ndarray
    light_cdf: np.ndarray
    light_total: float
    texels: np.ndarray

    def desc(self) -> SceneDesc:
        return SceneDesc(ptr(self.tris), len(self.tris), ptr(self.nodes), len(self.nodes), ptr(self.bvh_tri),
                         ptr(self.light_tri), ptr(self.light_cdf), len(self.light_tri), self.light_total,
                         ptr(self.texels), len(self.texels) // 3)


def material_rows(n: int, albedo=(1.0, 1.0, 1.0), emission=(0.0, 0.0, 0.0), metallic=0.0, ior=1.0, rough=0.5,
                  alpha=1.0, shader=0) -> np.ndarray:
    """(n, 11) material rows: ar ag ab er eg eb metallic ior rough alpha shader."""
    m = np.empty((n, 11), np.float64)
    m[:, 0:3] = albedo
    m[:, 3:6] = emission
    m[:, 6] = metallic
    m[:, 7] = ior
    m[:, 8] = rough
    m[:, 9] = alpha
    m[:, 10] = shader
    return m


class Scene:
    """raymotion::Scene in the product engine."""

    def __init__(self):
        self._h = lib().rm_scene_create()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().rm_scene_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def clear(self):
        lib().rm_scene_clear(self._h)

    def add_triangles(self, pos, mat, normals=None, uv=None) -> np.ndarray:
        """pos (n,3,3), mat (n,11) [, normals (n,3,3), uv (n,6)] -> triangle indices (-1 = degenerate)."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 9)
        n = len(pos)
        mat = np.ascontiguousarray(np.broadcast_to(mat, (n, 11)), np.float64)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(n, 9)
        uvs = None if uv is None else np.ascontiguousarray(uv, np.float64).reshape(n, 6)
        idx = np.empty(n, np.int32)
        rc = lib().rm_scene_add_triangles(self._h, n, ptr(pos), ptr(nrm), ptr(uvs), ptr(mat), ptr(idx))
        if rc:
            raise BackendError(rc, "rm_scene_add_triangles")
        return idx

    def set_positions(self, pos, normals=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 9)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(len(pos), 9)
        rc = lib().rm_scene_set_positions(self._h, len(pos), ptr(pos), ptr(nrm))
        if rc:
            raise BackendError(rc, "rm_scene_set_positions")

    def build_bvh(self, max_leaf_tris: int = 8):
        lib().rm_scene_build_bvh(self._h, max_leaf_tris)

    def refit_bvh(self):
        lib().rm_scene_refit_bvh(self._h)

    def counts(self) -> tuple[int, int, int]:
        a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        lib().rm_scene_counts(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def nodes(self):
        """Binary nodes in build order: boxes (n,6) f64, links (n,4) [left,right,offset,count], bvh_tri."""
        ntri, nnode, _ = self.counts()
        boxes = np.empty((nnode, 6), np.float64)
        links = np.empty((nnode, 4), np.int32)
        bvh_tri = np.empty(ntri if nnode else 0, np.int32)
        lib().rm_scene_nodes(self._h, ptr(boxes), ptr(links), ptr(bvh_tri))
        return boxes, links, bvh_tri

    def precompute(self):
        ntri, _, nlight = self.counts()
        n = np.empty((ntri, 3)); nn = np.empty((ntri, 3)); denom = np.empty(ntri); cdf = np.empty(nlight)
        total = ctypes.c_double()
        lib().rm_scene_precompute(self._h, ptr(n), ptr(nn), ptr(denom), ptr(cdf), ctypes.byref(total))
        return n, nn, denom, cdf, total.value

    def export(self) -> Exported:
        ntri, nnode, nlight = self.counts()
        tris = np.zeros(ntri, TRIANGLE_DTYPE)
        nodes = np.zeros(nnode, NODE_DTYPE)
        bvh_tri = np.zeros(ntri if nnode else 0, np.int32)
        light_tri = np.zeros(nlight, np.int32)
        cdf = np.zeros(nlight, np.float32)
        total = ctypes.c_float()
        lib().rm_scene_export(self._h, ptr(tris), ptr(nodes), ptr(bvh_tri), ptr(light_tri), ptr(cdf),
                              ctypes.byref(total))
        return Exported(tris, nodes, bvh_tri, light_tri, cdf, float(total.value), np.zeros(0, np.float32))


def camera_struct(pos, focal, m) -> Camera:
    """rm_camera from a position, focal length (pixels) and row-major 3x3 rotation."""
    m = np.asarray(m, np.float64).reshape(9)
    return Camera(float(pos[0]), float(pos[1]), float(pos[2]), float(focal), *[float(v) for v in m])


def default_frame(width: int, height: int, spp: int = 1, bounces: int = 5) -> Frame:
    """rm_frame with RenderConfig defaults after the environment switches."""
    f = Frame()
    lib().rm_default_frame(ctypes.byref(f), width, height, spp, bounces)
    return f


class Backend:
    """One rm_ctx on a HIP device."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().rm_create(ctypes.byref(h), device)
        if rc:
            raise BackendError(rc, lib().rm_last_error(None).decode())
        self._h = h.value
        # Counts scene uploads, so a caller can tell whether the scene it
        # uploaded is still the resident one.
        self.scene_serial = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().rm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            raise BackendError(rc, f"{what}: {lib().rm_last_error(self._h).decode()}")

    def set_scene(self, scene: Scene | Exported, gpu_tree: bool = False, rebuild: bool = False):
        """Uploads a scene. With gpu_tree the backend builds its traversal tree
        on the GPU (the scene needs no BVH; `rebuild` forces a new build,
        otherwise a resident GPU tree is refit while the triangle count stays)."""
        self.scene_serial += 1
        if isinstance(scene, Scene):
            if gpu_tree:
                self._check(lib().rm_scene_upload_gpu_tree(self._h, scene.handle, int(rebuild)),
                            "rm_scene_upload_gpu_tree")
            else:
                self._check(lib().rm_scene_upload(self._h, scene.handle), "rm_scene_upload")
        else:
            d = scene.desc()
            if gpu_tree:
                self._check(lib().rm_set_scene_gpu_tree(self._h, ctypes.byref(d), int(rebuild)),
                            "rm_set_scene_gpu_tree")
            else:
                self._check(lib().rm_set_scene(self._h, ctypes.byref(d)), "rm_set_scene")

    def render(self, frame: Frame, seed: int, want_hdr: bool = True):
        px = frame.width * frame.height
        rgb = np.empty(px * 3, np.uint8)
        hdr = np.empty(px * 3, np.float32) if want_hdr else None
        self._check(lib().rm_render(self._h, ctypes.byref(frame), seed, ptr(rgb), ptr(hdr)), "rm_render")
        return rgb, hdr

    def trace(self, origins, dirs, tmax=None, any_hit=False):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(o)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        tri = np.empty(n, np.int32)
        t = np.empty(n, np.float32); u = np.empty(n, np.float32); v = np.empty(n, np.float32)
        self._check(lib().rm_trace(self._h, ptr(o), ptr(d), ptr(tm), n, int(any_hit), ptr(tri), ptr(t), ptr(u),
                                   ptr(v)), "rm_trace")
        return tri, t, u, v

    def render_adjoint(self, frame: Frame, seed: int, d_hdr, tri_group, group_emission) -> np.ndarray:
        """Gradients (ngroup, 5) of sum(d_hdr * hdr) for the frame render(frame, seed) draws:
        albedo r, g, b, roughness and emission scale per group."""
        d = np.ascontiguousarray(d_hdr, np.float32).reshape(-1)
        if d.size != frame.width * frame.height * 3:
            raise ValueError("d_hdr must hold width * height * 3 values")
        tg = np.ascontiguousarray(tri_group, np.int32)
        ge = np.ascontiguousarray(group_emission, np.float32).reshape(-1, 3)
        grad = np.zeros((len(ge), 5), np.float64)
        self._check(lib().rm_render_adjoint(self._h, ctypes.byref(frame), seed, ptr(d), ptr(tg), ptr(ge), len(ge),
                                            ptr(grad)), "rm_render_adjoint")
        return grad

    def set_group_materials(self, tri_group, params, group_emission, light_cdf, light_total: float):
        """Rewrites the grouped triangles' albedo, roughness and emission
        (params (ngroup, 5): albedo r, g, b, roughness, emission scale) and the
        light CDF of the resident scene without re-sending it."""
        tg = np.ascontiguousarray(tri_group, np.int32)
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 5)
        ge = np.ascontiguousarray(group_emission, np.float32).reshape(-1, 3)
        cdf = np.ascontiguousarray(light_cdf, np.float32)
        if len(p) != len(ge):
            raise ValueError("params and group_emission need one row per group")
        self._check(lib().rm_set_group_materials(self._h, ptr(tg), tg.size, len(p), ptr(p), ptr(ge), ptr(cdf),
                                                 cdf.size, float(light_total)), "rm_set_group_materials")

    # ---- learned denoiser
    def kpn_set_weights(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        self._check(lib().rm_kpn_set_weights(self._h, ptr(p), p.size), "rm_kpn_set_weights")

    def kpn_get_weights(self) -> np.ndarray:
        out = np.zeros(lib().rm_kpn_param_count(), np.float32)
        self._check(lib().rm_kpn_get_weights(self._h, ptr(out), out.size), "rm_kpn_get_weights")
        return out

    def render_features(self, frame: Frame, seed: int):
        """(features (H, W, 16), noisy radiance (H, W, 3)) of a ReSTIR frame."""
        px = frame.width * frame.height
        feat = np.zeros((frame.height, frame.width, 16), np.float32)
        noisy = np.zeros((frame.height, frame.width, 3), np.float32)
        self._check(lib().rm_render_features(self._h, ctypes.byref(frame), seed, ptr(feat), ptr(noisy)),
                    "rm_render_features")
        assert feat.size == px * 16
        return feat, noisy

    def kpn_dataset_clear(self):
        self._check(lib().rm_kpn_dataset_clear(self._h), "rm_kpn_dataset_clear")

    def kpn_dataset_add(self, feat, noisy, reference):
        f = np.ascontiguousarray(feat, np.float32)
        h, w = f.shape[:2]
        n = np.ascontiguousarray(noisy, np.float32)
        r = np.ascontiguousarray(reference, np.float32)
        self._check(lib().rm_kpn_dataset_add(self._h, ptr(f), ptr(n), ptr(r), w, h), "rm_kpn_dataset_add")

    def kpn_train(self, steps: int, batch: int, lr: float, seed: int) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(lib().rm_kpn_train(self._h, steps, batch, lr, seed, ptr(losses)), "rm_kpn_train")
        return losses[:steps]

    def kpn_denoise(self, feat, noisy) -> np.ndarray:
        f = np.ascontiguousarray(feat, np.float32)
        h, w = f.shape[:2]
        out = np.zeros((h, w, 3), np.float32)
        self._check(lib().rm_kpn_denoise(self._h, ptr(f), ptr(np.ascontiguousarray(noisy, np.float32)), w, h,
                                         ptr(out)), "rm_kpn_denoise")
        return out

    def kpn_gradient(self, feat, noisy, reference):
        """(loss, gradient) of the current weights on images (N, H, W, C)."""
        f = np.ascontiguousarray(feat, np.float32)
        n, h, w = f.shape[:3]
        grad = np.zeros(lib().rm_kpn_param_count(), np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().rm_kpn_gradient(self._h, ptr(f), ptr(np.ascontiguousarray(noisy, np.float32)),
                                          ptr(np.ascontiguousarray(reference, np.float32)), n, w, h, ptr(grad),
                                          ptr(loss)), "rm_kpn_gradient")
        return float(loss[0]), grad

    def read_restir(self, width: int, height: int):
        px = width * height
        g = np.zeros(px, GBUFFER_DTYPE)
        ri = np.zeros(px, RESERVOIR_DTYPE)
        rm = np.zeros(px, RESERVOIR_DTYPE)
        shade = np.zeros(px * 3, np.float32)
        self._check(lib().rm_read_restir(self._h, ptr(g), ptr(ri), ptr(rm), ptr(shade)), "rm_read_restir")
        return g, ri, rm, shade

    def restir_merge(self, frame: Frame, seed: int, gbuf, rinit, hist_gbuf=None, hist_res=None,
                     prev_camera: Camera | None = None) -> np.ndarray:
        """The ReSTIR merge pass alone (restir_spatial) over given G-buffer and
        reservoirs (W*H records each); history arrays and camera optional."""
        px = frame.width * frame.height
        g = np.ascontiguousarray(gbuf, GBUFFER_DTYPE)
        ri = np.ascontiguousarray(rinit, RESERVOIR_DTYPE)
        hg = None if hist_gbuf is None else np.ascontiguousarray(hist_gbuf, GBUFFER_DTYPE)
        hr = None if hist_res is None else np.ascontiguousarray(hist_res, RESERVOIR_DTYPE)
        if g.size != px or ri.size != px or (hg is not None and hg.size != px) or (hr is not None and hr.size != px):
            raise ValueError("G-buffer and reservoir arrays must hold width * height records")
        out = np.zeros(px, RESERVOIR_DTYPE)
        cam = None if prev_camera is None else ctypes.byref(prev_camera)
        self._check(lib().rm_restir_merge(self._h, ctypes.byref(frame), seed, ptr(g), ptr(ri), ptr(hg), ptr(hr), cam,
                                          ptr(out)), "rm_restir_merge")
        return out

    def denoise_image(self, frame: Frame, color, gbuf, hist_color=None, hist_gbuf=None,
                      prev_camera: Camera | None = None) -> np.ndarray:
        """Temporal + frame.denoise_iterations a-trous passes over a linear image
        (W*H*3) and its G-buffer; history optional. Returns W*H*3 values."""
        px = frame.width * frame.height
        col = np.ascontiguousarray(color, np.float32).reshape(-1)
        g = np.ascontiguousarray(gbuf, GBUFFER_DTYPE)
        hc = None if hist_color is None else np.ascontiguousarray(hist_color, np.float32).reshape(-1)
        hg = None if hist_gbuf is None else np.ascontiguousarray(hist_gbuf, GBUFFER_DTYPE)
        if col.size != 3 * px or g.size != px or (hc is not None and hc.size != 3 * px) or \
                (hg is not None and hg.size != px):
            raise ValueError("colour must hold width * height * 3 values and the G-buffer width * height records")
        out = np.zeros(3 * px, np.float32)
        cam = None if prev_camera is None else ctypes.byref(prev_camera)
        self._check(lib().rm_denoise_image(self._h, ctypes.byref(frame), ptr(col), ptr(g), ptr(hc), ptr(hg), cam,
                                           ptr(out)), "rm_denoise_image")
        return out

    def read_tree(self):
        """(nodes (N, 32) float32, leaf-order triangles (T, 12) float32, stack depth)
        of the resident traversal tree, in the kernels' layout."""
        counts = np.zeros(3, np.int32)
        self._check(lib().rm_read_tree(self._h, None, 0, None, 0, ptr(counts)), "rm_read_tree")
        nodes = np.zeros((int(counts[0]), 32), np.float32)
        tris = np.zeros((int(counts[1]), 12), np.float32)
        self._check(lib().rm_read_tree(self._h, ptr(nodes), len(nodes), ptr(tris), len(tris), ptr(counts)),
                    "rm_read_tree")
        return nodes, tris, int(counts[2])

    def set_profiling(self, on: bool):
        self._check(lib().rm_set_profiling(self._h, int(on)), "rm_set_profiling")

    def set_schedule(self, schedule: int):
        """Path passes: 0 persistent lanes (default), 1 wavefront kernels with
        a persistent tail, 2 wavefront kernels only."""
        self._check(lib().rm_set_schedule(self._h, int(schedule)), "rm_set_schedule")

    def stats(self) -> Stats:
        s = Stats()
        self._check(lib().rm_get_stats(self._h, ctypes.byref(s)), "rm_get_stats")
        return s


__all__ = ["Scene", "Backend", "BackendError", "Exported", "material_rows", "camera_struct", "default_frame",
           "Frame", "Camera", "Stats", "_capi"]
