"""GPU path-replay adjoint (rm_render_adjoint) against the oracle and against
central finite differences, and its torch.autograd.Function wrapper.

* Against the oracle's FP32 adjoint (the same estimator and arithmetic): equal
  up to the order of the double accumulation.
* Against the oracle's FP64 adjoint on the Cornell box frame (256x256, 4 spp,
  depth 5): within 1e-3 of the largest gradient entry.
* Against central differences (h = 0.05, common random numbers) of the GPU
  forward on the 16-material inverse-rendering scene at 512x512, paired over
  96 seeds at 64 spp, on the unclamped light-sampling frame: within 2% where
  the differences resolve 2% (every albedo and emission entry above 20% of
  the largest), within 4 standard errors elsewhere.
* rm_set_group_materials renders and differentiates as a full upload.
"""
import math

import numpy as np
import pytest

import oracle_bindings as ob
from helpers import cornell
from raymotion_amd import Backend, scenes
from raymotion_amd.diff import DiffScene, Renderer, finite_differences, optimize, path_frame, render

pytestmark = pytest.mark.gpu

CORNELL_GROUPS = np.array([0] * 6 + [1] * 2 + [2] * 2 + [3] * 2 + [4] * 12 + [5] * 12, np.int32)


def _cornell_renderer(backend, size=256, spp=4, bounces=5):
    ds = DiffScene(cornell(rough=0.5).export(), CORNELL_GROUPS)
    f = path_frame(size, size, spp, bounces, scenes.cornell_camera())
    f.camera.focal = (size / 2) / math.tan(math.radians(20))
    return Renderer(backend, ds, f)


def test_adjoint_matches_oracle_cornell(backend):
    r = _cornell_renderer(backend)
    d = np.random.default_rng(1).uniform(0, 1, 256 * 256 * 3).astype(np.float32)
    g = r.adjoint(r.scene.initial, 4242, d)
    e = r.scene.exported(r.scene.initial)
    o32 = ob.render_adjoint(e, r.frame, 4242, d, r.scene.tri_group, r.scene.group_emission, precision=32)
    o64 = ob.render_adjoint(e, r.frame, 4242, d, r.scene.tri_group, r.scene.group_emission, precision=64)
    scale = np.abs(o64).max()
    print(f"cornell adjoint: max |gpu - oracle32| / max {np.abs(g - o32).max() / scale:.3g}, "
          f"max |gpu - oracle64| / max {np.abs(g - o64).max() / scale:.3g}")
    assert np.abs(g - o32).max() <= 1e-6 * scale
    assert np.abs(g - o64).max() <= 1e-3 * scale


@pytest.mark.parametrize("spp", [8, 32])
def test_adjoint_lane_streams(backend, spp):
    """spp 8 / 32 renders with 8 / 32 lane streams per pixel: the adjoint
    replays those streams."""
    r = _cornell_renderer(backend, size=64, spp=spp, bounces=4)
    d = np.random.default_rng(2).uniform(0, 1, 64 * 64 * 3).astype(np.float32)
    g = r.adjoint(r.scene.initial, 99, d)
    o32 = ob.render_adjoint(r.scene.exported(r.scene.initial), r.frame, 99, d, r.scene.tri_group,
                            r.scene.group_emission, precision=32)
    assert backend.stats().lanes == spp
    assert np.abs(g - o32).max() <= 1e-6 * np.abs(o32).max()


def _inverse_renderer(backend, size, spp, bounces=4):
    s, groups = scenes.inverse_scene(seed=0, n_spheres=24, subdivisions=2)
    ds = DiffScene(s.export(), groups)
    return Renderer(backend, ds, path_frame(size, size, spp, bounces, scenes.mixed_camera(0)))


def test_adjoint_matches_finite_differences(backend):
    """Central differences with common random numbers on the 16-material
    scene (327,698 triangles) at 512x512, depth 8, paired over 96 seeds of
    64 spp each, at an interior parameter point (albedo and roughness mapped
    into [0.05, 0.95]). The frame is the unclamped, light-sampling-only one
    (path_frame clamp=0, mis=False), for which the expected image is a
    polynomial in each albedo and linear in each emission scale and the
    replay's fixed decisions give its exact derivative. h = 0.05: the
    differences' noise comes from roulette and light-pick decisions that flip
    between the two sides (variance ~ 1/h), while the truncation error of the
    polynomial stays near 0.1% (the noise-free entries show it).

    Every entry above 10% of the largest is compared. An entry is informative
    where the paired difference's standard error is at most 0.5% of the
    gradient: each informative entry must agree to 2%, the others to 4
    standard errors. Every albedo and emission entry above 20% of the largest
    must be informative (roughness entries, whose GGX samples move
    visibility boundaries, and the weaker albedo entries need far more
    samples)."""
    s, groups = scenes.inverse_scene(seed=0)
    ds = DiffScene(s.export(), groups)
    f = path_frame(512, 512, 64, 8, scenes.mixed_camera(0), clamp=0.0, mis=False)
    r = Renderer(backend, ds, f)
    d = np.random.default_rng(3).uniform(0, 1, 512 * 512 * 3).astype(np.float32)
    p0 = ds.initial.copy()
    p0[:, :4] = 0.05 + 0.9 * p0[:, :4]
    mag = np.abs(r.adjoint(p0, 1, d))
    entries = [tuple(int(v) for v in ix) for ix in np.argwhere(mag >= 0.1 * mag.max())]
    rows, cols = np.array(entries).T
    K = 96
    G = np.array([r.adjoint(p0, 100 + k, d)[rows, cols] for k in range(K)])
    F = np.array([finite_differences(r, p0, 100 + k, d, entries, h=0.05) for k in range(K)])
    mg, mf = G.mean(axis=0), F.mean(axis=0)
    se = (F - G).std(axis=0, ddof=1) / np.sqrt(K)
    rel = np.abs(mg - mf) / np.abs(mg)
    informative = se <= 0.005 * np.abs(mg)
    main = (cols != 3) & (np.abs(mg) >= 0.2 * mag.max())
    for j, (gi, k) in enumerate(entries):
        print(f"group {gi} param {k}: |g|/max {abs(mg[j]) / mag.max():.3f} adjoint {mg[j]:+.5e} fd {mf[j]:+.5e} "
              f"rel {rel[j]:.4f} se/|g| {se[j] / abs(mg[j]):.4f}{'' if informative[j] else ' (4 se)'}")
    print(f"{len(entries)} entries >= 10% of the largest: {int(informative.sum())} informative (2% bound), "
          f"{int((~informative).sum())} at 4 standard errors; {int(main.sum())} main albedo / emission entries")
    assert np.all(rel[informative] <= 0.02)
    assert np.all(np.abs(mg - mf)[~informative] <= 4 * se[~informative])
    assert np.all(informative[main]) and main.sum() >= 8


def test_group_material_update_matches_full_upload(backend):
    """rm_set_group_materials (the Renderer's path after its first upload)
    renders and differentiates exactly as a full upload of the same
    parameters: identical images, and gradients equal up to the order of the
    adjoint's FP64 atomic accumulation (1e-9 of the largest entry), over
    parameter sets that move every group, the lights' emission and so the
    light CDF."""
    fast = _inverse_renderer(backend, 128, 16)
    full = Renderer(backend, fast.scene, fast.frame, full_upload=True)
    ds = fast.scene
    rng = np.random.default_rng(7)
    d = rng.uniform(0, 1, 128 * 128 * 3).astype(np.float32)
    sets = [ds.initial] + [ds.clamp(ds.initial * rng.uniform(0.6, 1.4, ds.initial.shape)) for _ in range(3)]
    want = [(full.forward(p, 50 + i), full.adjoint(p, 50 + i, d)) for i, p in enumerate(sets)]
    fast.forward(sets[-1], 1)  # the first call uploads the whole scene
    serial = backend.scene_serial
    for i, p in enumerate(sets):
        b, gb = fast.forward(p, 50 + i), fast.adjoint(p, 50 + i, d)
        a, ga = want[i]
        gd = float(np.abs(ga - gb).max() / np.abs(ga).max())
        print(f"set {i}: image equal {np.array_equal(a, b)}, gradient max diff / max {gd:.3g}")
        assert np.array_equal(a, b) and gd <= 1e-9
    assert backend.scene_serial == serial  # material updates only
    from raymotion_amd.scene import BackendError
    with pytest.raises(BackendError):
        backend.set_group_materials(ds.tri_group[:-1], sets[1], ds.group_emission, *ds.light_table(sets[1]))


def test_autograd_function(backend):
    import torch
    r = _cornell_renderer(backend, size=64, spp=4, bounces=4)
    p = torch.tensor(r.scene.initial, dtype=torch.float32, requires_grad=True)
    w = torch.rand(64, 64, 3, generator=torch.Generator().manual_seed(0))
    img = render(p, r, 11)
    assert img.shape == (64, 64, 3)
    (img * w).sum().backward()
    ref = r.adjoint(r.scene.initial.astype(np.float32).astype(np.float64), 11, w.numpy())
    np.testing.assert_allclose(p.grad.numpy(), ref.astype(np.float32), rtol=1e-6, atol=1e-6)


def _blocks(img, k=8):
    h, w, _ = img.shape
    return img[: h // k * k, : w // k * k].reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))


def test_inverse_rendering_recovers_parameters(backend):
    """The inverse-rendering configuration: the 16-material scene at
    512x512, depth 8, L2 at 16 spp against a 64 spp target, 200 Adam steps
    from parameters perturbed by up to 30% (seed 2), learning rate 3e-3 with
    cosine decay (at 1e-2 the steps wander along the valley where a light's
    emission scale trades against the albedos it illuminates). The loss
    falls, the block-averaged image approaches the target, and the strongly
    observed parameters move toward the ground truth."""
    s, groups = scenes.inverse_scene(seed=0)
    ds = DiffScene(s.export(), groups)
    cam = scenes.mixed_camera(0)
    r = Renderer(backend, ds, path_frame(512, 512, 16, 8, cam))
    hi = Renderer(backend, ds, path_frame(512, 512, 64, 8, cam))
    gt = ds.initial
    target = hi.forward(gt, 1)
    rng = np.random.default_rng(2)
    p0 = ds.clamp(gt * rng.uniform(0.7, 1.3, gt.shape))
    p, losses = optimize(r, target, p0, steps=200, lr=3e-3)

    def image_error(q):
        return float(np.mean((_blocks(hi.forward(q, 2)) - _blocks(target)) ** 2))

    e0, e1, floor = image_error(p0), image_error(p), image_error(gt)
    g = np.abs(r.adjoint(gt, 3, np.ones(512 * 512 * 3, np.float32)))
    strong = g >= 0.1 * g.max()
    err0, err = np.abs(p0 - gt)[strong].mean(), np.abs(p - gt)[strong].mean()
    print(f"loss {losses[:5].mean():.4g} -> {losses[-5:].mean():.4g}; block image error {e0:.4g} -> {e1:.4g} "
          f"(ground truth {floor:.4g}); {int(strong.sum())} strong parameters, mean |p - gt| {err0:.4f} -> {err:.4f}")
    assert losses[-5:].mean() < losses[:5].mean()
    assert e1 - floor < 0.5 * (e0 - floor)
    assert err < 0.7 * err0
