"""Differentiable rendering on the HIP backend.

Parameters are grouped per material: for each group, RGB albedo, roughness
and an emission scale (a group's emission is scale * its base emission). The
forward is the path-tracing frame of rm_render (legacy light sampling, no
ReSTIR, no adaptive stop); gradients come from rm_render_adjoint, the
path-replay adjoint that re-traces the forward's samples from their seeds
(memory O(pixels)). `render` exposes both through a torch.autograd.Function.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .scene import Backend, Exported, Frame

PARAM_NAMES = ("albedo_r", "albedo_g", "albedo_b", "roughness", "emission_scale")
NPARAM = len(PARAM_NAMES)


def path_frame(width: int, height: int, spp: int, bounces: int, camera, clamp: float = 16.0,
               mis: bool = True) -> Frame:
    """A frame that takes the path-tracing path the adjoint differentiates:
    one light sample per vertex, no spatial reuse, no adaptive stop.

    The adjoint holds the sample's discrete decisions (roulette, light pick,
    lobe choice) and its MIS weights fixed. That is the exact derivative of
    the expected image when the estimator is linear in those decisions:
    clamp=0 (no per-sample clamp) and mis=False (light sampling only). With
    the default per-sample clamp the clamped value of a path depends on the
    roulette outcome, and the reference's MIS weight for a BSDF-sampled light
    hit uses the previous vertex's one-candidate light estimate, not the
    light pdf, so the weights do not sum to one; the derivatives of those
    probabilities and weights are then left out of the gradient."""
    f = Frame()
    f.width, f.height, f.spp, f.bounces = width, height, spp, bounces
    f.resolution, f.clamp_value = 1.0, clamp
    f.adapt_min, f.adapt_step, f.adapt_rel, f.adapt_abs = 0, 8, 0.02, 0.002
    f.nomis = 0 if mis else 1
    f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius = 1, 0, 128, 16.0
    f.denoise, f.denoise_iterations, f.temporal = 0, 2, 0
    f.camera = camera
    return f


class DiffScene:
    """An exported scene whose materials are functions of per-group parameters.

    tri_group[i] is triangle i's group (-1: fixed). All triangles of a group
    must share albedo, roughness and emission; the group's base emission is
    that emission (the scale parameter starts at 1). A group without triangles
    keeps zero parameters and gets zero gradients.
    """

    def __init__(self, exported: Exported, tri_group, ngroup: int | None = None):
        self.src = exported
        self.tri_group = np.ascontiguousarray(tri_group, np.int32)
        if len(self.tri_group) != len(exported.tris):
            raise ValueError("tri_group needs one entry per triangle")
        self.ngroup = int(ngroup if ngroup is not None else self.tri_group.max() + 1)
        t = exported.tris
        self.group_emission = np.zeros((self.ngroup, 3), np.float32)
        self.initial = np.zeros((self.ngroup, NPARAM), np.float64)
        self.members = []
        for g in range(self.ngroup):
            idx = np.nonzero(self.tri_group == g)[0]
            self.members.append(idx)
            if len(idx) == 0:  # unused group: zero gradient, parameters unused
                continue
            for f in ("ar", "ag", "ab", "rough", "er", "eg", "eb"):
                if not np.all(t[f][idx] == t[f][idx[0]]):
                    raise ValueError(f"group {g}: triangles differ in {f}")
            i0 = idx[0]
            self.group_emission[g] = (t["er"][i0], t["eg"][i0], t["eb"][i0])
            self.initial[g] = (t["ar"][i0], t["ag"][i0], t["ab"][i0], t["rough"][i0], 1.0)
        # Light table: the CDF over the lights (area * max(luminance, 1e-6), the
        # weighting of Scene::add_triangle) is recomputed from the parameters.
        lt = exported.light_tri
        a, b, c = (t[k][lt].astype(np.float64) for k in ("v0", "v1", "v2"))
        self.light_area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)

    def clamp(self, params: np.ndarray) -> np.ndarray:
        """The parameter box the scene accepts: albedo in [0, 1], roughness in
        [0, 1], emission scale >= 1e-4 (a light stays a light)."""
        p = np.array(params, np.float64).reshape(self.ngroup, NPARAM)
        p[:, 0:4] = np.clip(p[:, 0:4], 0.0, 1.0)
        p[:, 4] = np.maximum(p[:, 4], 1e-4)
        return p

    def light_table(self, params):
        """(CDF float32, total) of the lights at these parameters."""
        p = self.clamp(params)
        lt = self.src.light_tri
        g = self.tri_group[lt]
        em = np.array([self.src.tris[k][lt] for k in ("er", "eg", "eb")], np.float32).T
        sel = g >= 0
        em[sel] = self.group_emission[g[sel]] * p[g[sel], 4:5].astype(np.float32)
        lum = 0.2126 * em[:, 0].astype(np.float64) + 0.7152 * em[:, 1] + 0.0722 * em[:, 2]
        w = self.light_area * np.maximum(lum, 1e-6)
        cdf = np.cumsum(w)
        total = float(cdf[-1]) if len(cdf) else 0.0
        return cdf.astype(np.float32), float(np.float32(total))

    def exported(self, params) -> Exported:
        p = self.clamp(params)
        tris = self.src.tris.copy()
        g = self.tri_group
        sel = g >= 0
        gs = g[sel]
        tris["ar"][sel] = p[gs, 0]
        tris["ag"][sel] = p[gs, 1]
        tris["ab"][sel] = p[gs, 2]
        tris["rough"][sel] = p[gs, 3]
        em = self.group_emission[gs] * p[gs, 4:5].astype(np.float32)
        tris["er"][sel] = em[:, 0]
        tris["eg"][sel] = em[:, 1]
        tris["eb"][sel] = em[:, 2]
        cdf, total = self.light_table(p)
        return Exported(tris, self.src.nodes, self.src.bvh_tri, self.src.light_tri, cdf, total, self.src.texels)


@dataclass
class Renderer:
    """Forward and adjoint renders of one DiffScene at a fixed frame.

    The first render uploads the whole scene; later ones, while that upload is
    still the backend's resident scene, rewrite only the grouped materials,
    the light emission and the light CDF (rm_set_group_materials), which
    renders exactly as a full upload of scene.exported(params) would.
    full_upload=True re-sends the whole scene every time."""
    backend: Backend
    scene: DiffScene
    frame: Frame
    full_upload: bool = False
    _serial: int = -1

    def load(self, params):
        """Makes the backend's resident scene scene.exported(params)."""
        if self.full_upload or self._serial != self.backend.scene_serial:
            self.backend.set_scene(self.scene.exported(params))
            self._serial = self.backend.scene_serial
            return
        p = self.scene.clamp(params)
        cdf, total = self.scene.light_table(p)
        self.backend.set_group_materials(self.scene.tri_group, p, self.scene.group_emission, cdf, total)

    def forward(self, params, seed: int) -> np.ndarray:
        self.load(params)
        _, hdr = self.backend.render(self.frame, seed, want_hdr=True)
        return hdr.reshape(self.frame.height, self.frame.width, 3)

    def adjoint(self, params, seed: int, d_hdr) -> np.ndarray:
        """(ngroup, 5) gradient of sum(d_hdr * forward(params, seed))."""
        self.load(params)
        return self.backend.render_adjoint(self.frame, seed, d_hdr, self.scene.tri_group, self.scene.group_emission)


def _torch_function():
    import torch

    class RenderFunction(torch.autograd.Function):
        """params (ngroup, 5) -> linear HDR image (H, W, 3); backward runs the
        HIP path-replay adjoint with the forward's seed."""

        @staticmethod
        def forward(ctx, params, renderer, seed, adjoint_seed):
            p = params.detach().to("cpu", torch.float64).numpy()
            img = renderer.forward(p, seed)
            ctx.renderer, ctx.p = renderer, p
            ctx.adjoint_seed = seed if adjoint_seed is None else adjoint_seed
            return torch.from_numpy(np.array(img)).to(params.device, torch.float32)

        @staticmethod
        def backward(ctx, grad_img):
            g = ctx.renderer.adjoint(ctx.p, ctx.adjoint_seed, grad_img.detach().to("cpu", torch.float32).numpy())
            return torch.from_numpy(g).to(grad_img.device, grad_img.dtype), None, None, None

    return RenderFunction


_RENDER_FN = None


def render(params, renderer: Renderer, seed: int, adjoint_seed: int | None = None):
    """Differentiable render: a torch tensor (ngroup, 5) of parameters to the
    (H, W, 3) float32 image, with gradients from the HIP adjoint. The backward
    pass replays the forward's own samples, or, with adjoint_seed, an
    independent set (an unbiased gradient of a loss that is nonlinear in the
    image, such as L2, needs the image and its derivative decorrelated)."""
    global _RENDER_FN
    if _RENDER_FN is None:
        _RENDER_FN = _torch_function()
    return _RENDER_FN.apply(params, renderer, seed, adjoint_seed)


def finite_differences(renderer: Renderer, params, seed: int, d_hdr, entries, h: float = 1e-3) -> np.ndarray:
    """Central differences of sum(d_hdr * image) with common random numbers
    (the same seed on both sides) for the listed (group, param) entries."""
    p0 = np.array(params, np.float64).reshape(renderer.scene.ngroup, NPARAM)
    d = np.asarray(d_hdr, np.float64).reshape(renderer.frame.height, renderer.frame.width, 3)
    out = []
    for g, k in entries:
        hi, lo = p0.copy(), p0.copy()
        hi[g, k] += h
        lo[g, k] -= h
        fp = float(np.sum(d * renderer.forward(hi, seed)))
        fm = float(np.sum(d * renderer.forward(lo, seed)))
        out.append((fp - fm) / (2.0 * h))
    return np.array(out)


def optimize(renderer: Renderer, target, p0, steps: int = 200, lr: float = 1e-2, seed0: int = 1000,
             log_every: int = 0, final_lr_fraction: float = 0.1):
    """Adam on the pixel L2 loss against `target` (H, W, 3) with new seeds per
    iteration (forward and adjoint decorrelated); the learning rate follows a
    cosine from lr to final_lr_fraction * lr (1.0: constant). Returns (final
    params, per-step losses)."""
    import torch
    tgt = torch.as_tensor(np.asarray(target, np.float32))
    p = torch.tensor(np.asarray(p0, np.float64), dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([p], lr=lr)
    losses = []
    for it in range(steps):
        c = 0.5 * (1.0 + np.cos(np.pi * it / max(steps - 1, 1)))
        for grp in opt.param_groups:
            grp["lr"] = lr * (final_lr_fraction + (1.0 - final_lr_fraction) * c)
        opt.zero_grad()
        img = render(p, renderer, seed0 + 2 * it, seed0 + 2 * it + 1)
        loss = torch.mean((img - tgt) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            p.copy_(torch.as_tensor(renderer.scene.clamp(p.detach().numpy()), dtype=torch.float32))
        losses.append(float(loss.detach()))
        if log_every and (it % log_every == 0 or it == steps - 1):
            print(f"iter {it}: loss {losses[-1]:.6g}", flush=True)
    return p.detach().numpy().astype(np.float64), np.array(losses)


__all__ = ["PARAM_NAMES", "NPARAM", "DiffScene", "Renderer", "path_frame", "render", "finite_differences",
           "optimize"]
