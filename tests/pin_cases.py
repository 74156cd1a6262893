"""Transport and filter pins as functions of a runner, so that the same checks
run against the CPU oracle and against the HIP backend (tests/test_pins_oracle.py,
tests/test_gpu_pins.py). Expectations come from transport_pins.py.

A runner provides
  hdr_batches(exported, frame, seeds) -> (B, W*H, 3) linear images, one per seed
  gbuffer(exported, frame, seed)      -> W*H G-buffer records of a ReSTIR frame
  merge(exported, frame, seed, gbuf, rinit, hist_gbuf, hist_res, prev_camera) -> W*H reservoirs
  denoise(frame, colour, gbuf, hist_colour, hist_gbuf, prev_camera) -> W*H*3 values
  scale: sample-count multiplier (more samples where the runner is fast)
  batches: independent seeds per Monte Carlo check
"""
from __future__ import annotations

import math

import numpy as np

import transport_pins as tp
from helpers import frame
from raymotion_amd._capi import GBUFFER_DTYPE, RESERVOIR_DTYPE

_DL = None
_FURNACE = None
_SLAB = None


def direct_light():
    global _DL
    if _DL is None:
        _DL = tp.DirectLight()
    return _DL


def furnace():
    global _FURNACE
    if _FURNACE is None:
        _FURNACE = tp.FurnaceSphere()
    return _FURNACE


def slab():
    global _SLAB
    if _SLAB is None:
        _SLAB = tp.GlassSlab()
    return _SLAB


# ---------------------------------------------------------------- transport
def direct_light_case(runner, *, depth=1, nomis=1, candidates=1, restir=False, size=16, spp=256, batches=None):
    """NEE (candidates 1), RIS (candidates > 1) or the ReSTIR frame's primary
    direct light against the quadrature of the direct-lighting integral."""
    sc = direct_light()
    cam = sc.camera(size)
    f = tp.pin_frame(size, size, spp * runner.scale, depth, cam, nomis=nomis, candidates=candidates, restir=restir)
    got = runner.hdr_batches(sc.exported, f, tp.seeds(batches or runner.batches))
    if depth == 1:
        want = sc.expectation_depth1(cam, size, size, nomis=bool(nomis))
    else:
        assert depth == 2 and not nomis
        want = sc.expectation_depth2_mis(cam, size, size)
    return tp.check_mc(got, want, what=f"direct light depth {depth} nomis {nomis} candidates {candidates}")


def furnace_case(runner, depth, *, restir=False, candidates=1, size=12, spp=256, batches=None):
    """Multi-bounce estimate (NEE at every vertex, roulette after the second)
    in the emitting sphere against E * sum_k T^k 1. The sphere is an
    icosphere (81,920 facets with smooth normals): a 3e-4 relative allowance
    covers its difference from the exact sphere."""
    sc = furnace()
    cam = sc.camera(size)
    f = tp.pin_frame(size, size, spp * runner.scale, depth, cam, nomis=1, candidates=candidates, restir=restir)
    got = runner.hdr_batches(sc.exported, f, tp.seeds(batches or runner.batches))
    want = sc.expectation(cam, size, size, depth)
    return tp.check_mc(got, want, rel_sys=3e-4, what=f"furnace depth {depth} restir {restir}")


def slab_case(runner, depth, *, restir=False, size=12, spp=256, batches=None):
    """Smooth dielectric transport through a glass slab between emitting walls."""
    sc = slab()
    cam = sc.camera(size)
    f = tp.pin_frame(size, size, spp * runner.scale, depth, cam, nomis=1, candidates=1, restir=restir)
    got = runner.hdr_batches(sc.exported, f, tp.seeds(batches or runner.batches))
    want = sc.expectation(cam, size, size, depth)
    return tp.check_mc(got, want, what=f"glass slab depth {depth} restir {restir}")


def hue_case(runner, size=32):
    """shader == 1: the G-buffer albedo of every hit is hsv2rgb(h, 0.8, 1),
    which the reference returns in 0..255."""
    for az in (0.0, 120.0, 240.0):
        exported, cam = tp.hue_scene(az)
        cam.focal = 40.0 * size / 32
        f = frame(size, size, 1, 2, cam=cam, candidates=4, spatial=1, radius=0.1, denoise=0, temporal=0)
        g = runner.gbuffer(exported, f, 777)
        assert np.all(g["kind"] == 1)
        want = tp.hue_expected(g["pos"])
        err = float(np.abs(g["albedo"] / want - 1).max())
        assert err < 2e-6, (az, err)


# ---------------------------------------------------------------- denoiser
def _denoise_frame(w, h, iterations, cam=None):
    from helpers import default_camera
    f = frame(w, h, 1, 1, cam=cam if cam is not None else default_camera(w), denoise=1, iterations=iterations)
    return f


def atrous_box_case(runner, w=24, h=20):
    """Flat G-buffer and equal-luminance colours: every edge-stopping weight is
    one, so each pass is the plain mean over the in-image, non-sky 5x5 taps at
    stride 2^step; sky pixels give zero."""
    rng = np.random.default_rng(3)
    g = tp.gbuffer(w, h)
    sky = rng.random(w * h) < 0.1
    g["kind"][sky] = 0
    col = tp.constant_luma_colours(w * h, rng)
    for iters in (1, 2, 3):
        out = runner.denoise(_denoise_frame(w, h, iters), col.reshape(-1), g, None, None, None).reshape(h, w, 3)
        want = col.reshape(h, w, 3).copy()
        valid = ~sky.reshape(h, w)
        for s in range(iters):
            want = tp.box_filter(want, valid, s)
        err = np.abs(out - want).max()
        assert err < 1e-5, (iters, err)


def _two_region_case(runner, w, h, field, values, weight_of, rtol=1e-5):
    """Left half and right half differ in one G-buffer field or in colour
    luminance; the pixel just left of the edge averages its 5x5 taps with the
    right-half taps weighted by weight_of (a closed form)."""
    g = tp.gbuffer(w, h)
    col = np.ones((w * h, 3))
    left = (np.arange(w * h) % w) < w // 2
    if field == "colour":
        col[left] = values[0]
        col[~left] = values[1]
    else:
        g[field][left] = values[0]
        g[field][~left] = values[1]
    out = runner.denoise(_denoise_frame(w, h, 1), col.reshape(-1), g, None, None, None).reshape(h, w, 3)
    y = h // 2
    for x, own, other in ((w // 2 - 1, 0, 1), (w // 2 - 2, 0, 1), (w // 2, 1, 0)):
        n_same = 5 * sum(1 for dx in range(-2, 3) if (x + dx < w // 2) == (own == 0))
        n_other = 25 - n_same
        wt = weight_of(own, other)
        c_own = col[y * w + x]
        c_oth = col[y * w + (w // 2 if own == 0 else w // 2 - 1)]
        want = (n_same * c_own + n_other * wt * c_oth) / (n_same + n_other * wt)
        assert np.allclose(out[y, x], want, rtol=rtol, atol=1e-7), (field, x, out[y, x], want)


def atrous_weights_case(runner, w=16, h=12):
    from helpers import default_camera
    cam = default_camera(w)
    p = (np.array([0.0, 0.0, 5.0]), np.array([0.3, 0.0, 5.0]))
    dc = [float(np.linalg.norm(v - np.array([cam.x, cam.y, cam.z]))) for v in p]
    # position step of 0.3 along x: exp(-d / (0.1 dc + 1e-4)), dc of the centre pixel
    _two_region_case(runner, w, h, "pos", p, lambda a, b: math.exp(-0.3 / (0.1 * dc[a] + 1e-4)))
    # normal step: max(n . n0, 0)^16
    n1 = np.array([math.sin(0.4), 0.0, math.cos(0.4)])
    _two_region_case(runner, w, h, "normal", ((0.0, 0.0, 1.0), n1), lambda a, b: math.cos(0.4) ** 16)
    # albedo step: exp(-|da| / 0.25)
    _two_region_case(runner, w, h, "albedo", ((0.5, 0.5, 0.5), (0.6, 0.5, 0.3)),
                     lambda a, b: math.exp(-math.sqrt(0.1 ** 2 + 0.2 ** 2) / 0.25))
    # luminance step: exp(-|l - l0| / (0.3 max(l0, 0.5) + 1e-3)), l0 the centre pixel's
    lums = (0.8, 1.4)
    _two_region_case(runner, w, h, "colour", (np.full(3, lums[0]), np.full(3, lums[1])),
                     lambda a, b: math.exp(-abs(lums[b] - lums[a]) / (0.3 * max(lums[a], 0.5) + 1e-3)))


def atrous_fixed_point_case(runner, w=20, h=14):
    """A constant image over a constant G-buffer is a fixed point."""
    g = tp.gbuffer(w, h)
    col = np.tile([0.7, 0.2, 1.3], w * h)
    out = runner.denoise(_denoise_frame(w, h, 4), col, g, None, None, None)
    assert np.allclose(out, col, rtol=1e-6, atol=0)


def _reprojecting_gbuffer(cam, w, h, depth=4.0):
    """Positions on the pixel-centre rays at the given distance, so that a
    static camera reprojects every pixel onto itself."""
    y, x = np.mgrid[0:h, 0:w]
    sx = (x + 0.5 - w * 0.5).reshape(-1)
    sy = (h * 0.5 - (y + 0.5)).reshape(-1)
    t = np.stack([sx, sy, np.full(sx.shape, cam.focal)], 1)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    m = np.array([cam.m0, cam.m1, cam.m2, cam.m3, cam.m4, cam.m5, cam.m6, cam.m7, cam.m8]).reshape(3, 3)
    pos = np.array([cam.x, cam.y, cam.z]) + depth * (t @ m.T)
    g = tp.gbuffer(w, h)
    g["pos"] = pos
    g["normal"] = -(t @ m.T)
    return g


def temporal_case(runner, w=16, h=12):
    """Static camera with a known history: 0.2 cur + 0.8 clamp(h, 3x3 min/max
    of cur); history rejected beyond 0.02 dc + 1e-3 in position or at a normal
    dot of 0.9 or below; dielectric (kind 3) pixels pass through."""
    from helpers import default_camera
    cam = default_camera(w)
    g = _reprojecting_gbuffer(cam, w, h)
    rng = np.random.default_rng(5)
    cur = rng.uniform(0.2, 2.0, (w * h, 3))
    hist = rng.uniform(0.0, 3.0, (w * h, 3))
    hg = g.copy()
    dc = np.linalg.norm(g["pos"] - np.array([cam.x, cam.y, cam.z]), axis=1)
    accept = np.ones(w * h, bool)
    # position offsets just inside / just outside the threshold
    idx = rng.permutation(w * h)
    inside, outside, n_ok, n_bad, dielectric = idx[:20], idx[20:40], idx[40:60], idx[60:80], idx[80:90]
    thr = 0.02 * dc + 1e-3
    for sel, k in ((inside, 0.9), (outside, 1.1)):
        hg["pos"][sel] += (k * thr[sel])[:, None] * _perp(g["normal"][sel])
    accept[outside] = False
    for sel, dot in ((n_ok, 0.905), (n_bad, 0.895)):
        hg["normal"][sel] = _tilt(g["normal"][sel], math.acos(dot))
    accept[n_bad] = False
    g["kind"][dielectric] = 3
    accept[dielectric] = False
    out = runner.denoise(_denoise_frame(w, h, 0), cur.reshape(-1), g, hist.reshape(-1), hg, cam).reshape(-1, 3)
    c3 = cur.reshape(h, w, 3)
    pad_lo = np.pad(c3, ((1, 1), (1, 1), (0, 0)), constant_values=np.inf)
    pad_hi = np.pad(c3, ((1, 1), (1, 1), (0, 0)), constant_values=-np.inf)
    mn = np.min([pad_lo[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)], axis=0).reshape(-1, 3)
    mx = np.max([pad_hi[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)], axis=0).reshape(-1, 3)
    hc = np.minimum(np.maximum(hist, mn), mx)
    want = np.where(accept[:, None], hc + (cur - hc) * 0.2, cur)
    assert np.allclose(out, want, rtol=1e-6, atol=1e-7), float(np.abs(out - want).max())


def _perp(n):
    a = np.cross(n, [0.0, 1.0, 0.0])
    return a / np.linalg.norm(a, axis=1, keepdims=True)


def _tilt(n, angle):
    p = _perp(n)
    return n * math.cos(angle) + p * math.sin(angle)


# ---------------------------------------------------------------- ReSTIR merge
def _merge_setup(w, h):
    """Every pixel sees the same shading point (plane normal +z) and holds the
    same light sample; stored p_hat, sum weights and M differ per pixel."""
    sc = direct_light()
    g = tp.gbuffer(w, h, pos=(0.2, 0.1, 0.0), normal=(0.0, 0.0, 1.0), albedo=sc.albedo)
    wo = np.array([-0.5, -0.4, 0.7])
    g["wo"] = wo / np.linalg.norm(wo)
    g["rough"] = sc.rough
    g["metallic"] = sc.metallic
    rng = np.random.default_rng(11)
    r = np.zeros(w * h, RESERVOIR_DTYPE)
    # a sample point off every surface of the scene, so that no shadow ray
    # (from points within 1.5 of the origin on z = 0) is blocked
    r["light_pos"] = (-2.5, -2.0, 2.0)
    r["light_normal"] = (0.0, 0.0, -1.0)
    r["emission"] = sc.emission
    r["alpha"] = 1.0
    r["p_hat"] = rng.uniform(0.5, 2.0, w * h)
    r["sum_ws"] = rng.uniform(0.1, 1.0, w * h)
    r["sum_wu"] = r["sum_ws"] * rng.uniform(1.0, 1.5, w * h)
    r["m"] = rng.integers(1, 9, w * h)
    r["valid"] = 1
    return sc, g, r


def target_of(g0, res0):
    """p_hat of a reservoir's light sample at a G-buffer point."""
    p = np.asarray(g0["pos"], np.float64)
    lp = np.asarray(res0["light_pos"], np.float64)
    d = lp - p
    dist = np.linalg.norm(d)
    wi = d / dist
    n = np.asarray(g0["normal"], np.float64)
    f, _ = tp.opaque_bsdf(wi, np.asarray(g0["wo"], np.float64), n, float(g0["rough"]), float(g0["metallic"]),
                          np.asarray(g0["albedo"], np.float64))
    scos = float(np.dot(n, wi))
    return float((np.asarray(res0["emission"]) * f * scos * float(res0["alpha"])) @ tp.LUM)


def merge_case(runner, w=12, h=10, spatial=4, radius=3.0, mcap=20):
    """Identical light samples: each merged neighbour adds p_hat_cur * sum_w /
    p_hat_stored to both weight sums and its M (capped), whichever sample is
    selected; neighbours come from the stream's offsets; a neighbour normal at
    a dot below 0.9 is skipped."""
    sc, g, r = _merge_setup(w, h)
    tilted = (np.arange(w * h) % 3) == 2
    for dot in (0.905, 0.895):
        gg = g.copy()
        gg["normal"][tilted] = _tilt(np.broadcast_to([0.0, 0.0, 1.0], (int(tilted.sum()), 3)), math.acos(dot))
        f = frame(w, h, 1, 2, cam=tp.DirectLight.camera(w), candidates=4, spatial=spatial, radius=radius, mcap=mcap)
        out = runner.merge(sc.exported, f, 4242, gg, r, None, None, None)
        ok_normal = lambda pid, nid: (not tilted[nid] and not tilted[pid]) or \
            (float(np.dot(gg["normal"][pid], gg["normal"][nid])) >= 0.9)
        nbrs = tp.spatial_neighbours(4242, w, h, spatial, radius, ok_normal)
        for pid in range(w * h):
            p_cur = target_of(gg[pid], r[pid])
            su = float(r["sum_wu"][pid]) + sum(p_cur * float(r["sum_wu"][n]) / float(r["p_hat"][n]) for n in nbrs[pid])
            ss = float(r["sum_ws"][pid]) + sum(p_cur * float(r["sum_ws"][n]) / float(r["p_hat"][n]) for n in nbrs[pid])
            m = int(r["m"][pid])
            for n in nbrs[pid]:
                m = min(m + int(r["m"][n]), mcap)
            o = out[pid]
            assert o["valid"] == 1 and int(o["m"]) == m, (dot, pid, o["m"], m)
            assert abs(float(o["sum_wu"]) - su) <= 2e-5 * su, (dot, pid, o["sum_wu"], su)
            assert abs(float(o["sum_ws"]) - ss) <= 2e-5 * ss, (dot, pid, o["sum_ws"], ss)
            # the selected sample keeps its stored p_hat, or takes the target
            # at this pixel when a neighbour's copy was selected
            own = float(o["p_hat"]) == float(r["p_hat"][pid])
            assert own or (nbrs[pid] and abs(float(o["p_hat"]) - p_cur) <= 2e-5 * p_cur), (pid, o["p_hat"], p_cur)


def merge_temporal_case(runner, w=12, h=10, mcap=20):
    """Temporal reuse only (no spatial neighbours): the history reservoir at
    the reprojected pixel adds p_hat_cur * hsum / hp_hat when the history
    point lies within 0.02 dc + 1e-3 and its normal at a dot above 0.9; M is
    capped at mcap."""
    sc, g, r = _merge_setup(w, h)
    cam = tp.look_at((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 20.0, up=(0.0, 1.0, 0.0))
    g = _merge_gbuffer_on_rays(g, cam, w, h)
    rng = np.random.default_rng(13)
    hr = r.copy()
    hr["p_hat"] = rng.uniform(0.5, 2.0, w * h)
    hr["m"] = rng.integers(5, 20, w * h)
    hg = g.copy()
    dc = np.linalg.norm(g["pos"] - np.array([cam.x, cam.y, cam.z]), axis=1)
    thr = 0.02 * dc + 1e-3
    idx = rng.permutation(w * h)
    far, bent = idx[:15], idx[15:30]
    hg["pos"][far] += (1.1 * thr[far])[:, None] * np.array([1.0, 0.0, 0.0])
    near = idx[30:45]
    hg["pos"][near] += (0.9 * thr[near])[:, None] * np.array([1.0, 0.0, 0.0])
    hg["normal"][bent] = _tilt(np.broadcast_to([0.0, 0.0, 1.0], (15, 3)), math.acos(0.895))
    accept = np.ones(w * h, bool)
    accept[far] = False
    accept[bent] = False
    f = frame(w, h, 1, 2, cam=cam, candidates=4, spatial=1, radius=0.1, mcap=mcap)
    out = runner.merge(sc.exported, f, 99, g, r, hg, hr, cam)
    for pid in range(w * h):
        p_cur = target_of(g[pid], r[pid])
        su, ss, m = float(r["sum_wu"][pid]), float(r["sum_ws"][pid]), int(r["m"][pid])
        if accept[pid]:
            su += p_cur * float(hr["sum_wu"][pid]) / float(hr["p_hat"][pid])
            ss += p_cur * float(hr["sum_ws"][pid]) / float(hr["p_hat"][pid])
            m = min(m + int(hr["m"][pid]), mcap)
        o = out[pid]
        assert int(o["m"]) == m, (pid, o["m"], m, accept[pid])
        assert abs(float(o["sum_wu"]) - su) <= 2e-5 * su, (pid, o["sum_wu"], su)
        assert abs(float(o["sum_ws"]) - ss) <= 2e-5 * ss, (pid, o["sum_ws"], ss)


def _merge_gbuffer_on_rays(g, cam, w, h):
    """Pixel-centre hits of a static camera on the plane z = 0, so that every
    pixel reprojects onto itself."""
    y, x = np.mgrid[0:h, 0:w]
    sx = (x + 0.5 - w * 0.5).reshape(-1)
    sy = (h * 0.5 - (y + 0.5)).reshape(-1)
    t = np.stack([sx, sy, np.full(sx.shape, cam.focal)], 1)
    m = np.array([cam.m0, cam.m1, cam.m2, cam.m3, cam.m4, cam.m5, cam.m6, cam.m7, cam.m8]).reshape(3, 3)
    d = t @ m.T
    o = np.array([cam.x, cam.y, cam.z])
    pos = o - (o[2] / d[:, 2])[:, None] * d
    g = g.copy()
    g["pos"] = pos
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    g["wo"] = -dn
    return g


ALL_FILTER_CASES = [atrous_box_case, atrous_weights_case, atrous_fixed_point_case, temporal_case]
ALL_MERGE_CASES = [merge_case, merge_temporal_case]
