"""Pins the oracle's ReSTIR DI passes (restir_generate, restir_spatial and
restir_shade, pathtrace.metal:817-1258) on the CPU beyond restating them.

* Unbiasedness without reuse: with the spatial pass reduced to the identity
  (a radius under half a pixel, so every neighbour rounds to the pixel itself
  and is skipped, pathtrace.metal:960-963) and without MIS, the ReSTIR frame
  estimates the same integral as the legacy next-event-estimation megakernel
  (render_paths, one light candidate): per-block and per-pixel means over many
  seeds agree within their standard errors. This checks the RIS weights
  W = sum_ws / (M p_hat) of the initial reservoirs (with visibility in sum_ws),
  the reservoir selection, and the shade pass's use of them, against an
  independent estimator. (Spatial reuse with the reference's 1/M combination
  is biased by construction, and so is MIS with the reservoir's approximate
  pdf, so neither is compared here.)
* Merge invariants of the spatial and temporal pass: M never exceeds the cap
  and only grows, the sums only grow, the visible sum never exceeds the
  unoccluded one, and a merged reservoir holds a light sample drawn by some
  pixel of this frame or of the history.
"""
import numpy as np

import oracle_bindings as ob
from helpers import cornell, cornell_frame

W = H = 32
SEEDS = 800


def _mean_images(exported, f, first_seed):
    acc = np.zeros((H, W, 3))
    acc2 = np.zeros((H, W, 3))
    for s in range(SEEDS):
        _, hdr, info = ob.render_frame(exported, f, first_seed + s)
        img = hdr.reshape(H, W, 3).astype(np.float64)
        acc += img
        acc2 += img * img
    return info, acc / SEEDS, acc2 / SEEDS


def test_restir_without_reuse_is_unbiased_against_nee():
    e = cornell().export()
    common = dict(denoise=0, temporal=0, clamp=1e9, nomis=1)
    f_restir = cornell_frame(W, H, 4, 5, candidates=8, spatial=1, radius=0.1, **common)
    f_nee = cornell_frame(W, H, 4, 5, candidates=1, spatial=0, **common)
    # the spatial pass is the identity at this radius
    _, _, dbg = ob.render_frame(e, f_restir, 7, debug=True)
    assert dbg["path"] == 1
    for k in dbg["rinit"].dtype.names:
        assert np.array_equal(dbg["rinit"][k], dbg["rmerged"][k]), k
    info_a, mean_a, sq_a = _mean_images(e, f_restir, 1000)
    info_b, mean_b, sq_b = _mean_images(e, f_nee, 50000)
    assert info_a["path"] == 1 and info_b["path"] == 0
    var_a = np.maximum(sq_a - mean_a ** 2, 0) / (SEEDS - 1)
    var_b = np.maximum(sq_b - mean_b ** 2, 0) / (SEEDS - 1)
    # per pixel
    z = (mean_a - mean_b) / np.sqrt(var_a + var_b + 1e-30)
    z_rms = float(np.sqrt(np.mean(z ** 2)))
    print(f"pixel z rms {z_rms:.3f}, max {np.abs(z).max():.2f}")
    assert 0.7 < z_rms < 1.25
    # per 8x8 block (block means of independent pixels)
    k = 8
    bm_a = mean_a.reshape(H // k, k, W // k, k, 3).mean(axis=(1, 3))
    bm_b = mean_b.reshape(H // k, k, W // k, k, 3).mean(axis=(1, 3))
    bv_a = var_a.reshape(H // k, k, W // k, k, 3).sum(axis=(1, 3)) / k ** 4
    bv_b = var_b.reshape(H // k, k, W // k, k, 3).sum(axis=(1, 3)) / k ** 4
    zb = (bm_a - bm_b) / np.sqrt(bv_a + bv_b + 1e-30)
    print(f"block z rms {np.sqrt(np.mean(zb ** 2)):.3f}, max {np.abs(zb).max():.2f}; "
          f"means {mean_a.mean():.5f} vs {mean_b.mean():.5f}")
    assert np.sqrt(np.mean(zb ** 2)) < 2.0
    assert np.abs(zb).max() < 4.5
    assert abs(mean_a.mean() / mean_b.mean() - 1) < 0.005


def test_spatial_and_temporal_merge_invariants():
    e = cornell().export()
    f = cornell_frame(48, 48, 2, 4, candidates=8, spatial=3, radius=10.0, mcap=20, denoise=0, temporal=1)
    hist = ob.History()
    prev_lights = None
    for frame_no in range(3):
        _, _, dbg = ob.render_frame(e, f, 12345 + frame_no, history=hist, debug=True)
        g, ri, rm = dbg["gbuf"], dbg["rinit"], dbg["rmerged"]
        assert dbg["has_history"] == (1 if frame_no > 0 else 0)
        act = (g["kind"] == 1) & (ri["valid"] != 0)
        assert act.sum() > 1000
        assert np.all(rm["m"][act] <= 20)
        assert np.all(rm["m"][act] >= np.minimum(ri["m"][act], 20))
        assert np.all(rm["sum_wu"][act] >= ri["sum_wu"][act])
        assert np.all(rm["sum_ws"][act] >= ri["sum_ws"][act])
        assert np.all(rm["sum_ws"][act] <= rm["sum_wu"][act] * (1 + 1e-5))
        grew = act & (rm["m"] > ri["m"])
        assert grew.mean() > 0.5 * act.mean()
        # every merged sample was drawn by some pixel, now or in the history
        pool = {tuple(p) for p in ri["light_pos"][act]}
        if prev_lights is not None:
            pool |= prev_lights
        assert all(tuple(p) in pool for p in rm["light_pos"][act & (rm["sum_ws"] > 0)])
        prev_lights = {tuple(p) for p in rm["light_pos"][act]}
        # pixels that are not opaque hits pass through unchanged
        other = ~act
        for k in ("sum_ws", "sum_wu", "m", "p_hat"):
            assert np.array_equal(rm[k][other], ri[k][other])
