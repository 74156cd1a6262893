"""The oracle's frame renderer on a 4-wide tree in the backend's layout
(oracle_bindings.wide_tree, used by the full-size GPU comparisons with the
backend's own tree): on a tree built here (one node, four leaves) it finds
the brute-force hits, and the Cornell frame equals the one rendered on the
scene's binary tree bit for bit (no ray grazes a box edge there)."""
import numpy as np

import oracle_bindings as ob
from helpers import cornell, cornell_frame
from raymotion_amd.rng import random_rays


def _one_node_tree(e):
    """Root with four leaf children over consecutive quarters of the triangles."""
    v = np.stack([e.tris["v0"], e.tris["v1"], e.tris["v2"]], axis=1).astype(np.float32)  # (n, 3, 3)
    n = len(v)
    node = np.zeros(32, np.float32)
    node[0:24] = np.tile(np.array([np.inf, -np.inf], np.float32).repeat(4), 3)
    node[24:32].view(np.int32)[:] = [0, 0, 0, 0, -1, -1, -1, -1]
    cuts = np.linspace(0, n, 5).astype(int)
    for c in range(4):
        a, b = cuts[c], cuts[c + 1]
        if b <= a:
            continue
        lo, hi = v[a:b].reshape(-1, 3).min(axis=0), v[a:b].reshape(-1, 3).max(axis=0)
        for ax in range(3):
            node[8 * ax + c] = lo[ax]
            node[8 * ax + 4 + c] = hi[ax]
        node[24:28].view(np.int32)[c] = a
        node[28:32].view(np.int32)[c] = b - a
    geo = np.zeros((n, 12), np.float32)
    geo[:, 0:3], geo[:, 3:6], geo[:, 6:9] = v[:, 0], v[:, 1], v[:, 2]
    geo[:, 9].view(np.int32)[:] = np.arange(n, dtype=np.int32)
    return node.reshape(1, 32), geo


def test_wide_tree_frames_equal_binary_tree():
    s = cornell()
    e = s.export()
    nodes, geo = _one_node_tree(e)
    o, d = random_rays(1 << 14, 3, lo=(-0.5, -0.5, -0.5), hi=(1.5, 1.5, 1.5))
    w_tri, w_t, _ = ob.trace_wide(nodes, geo, o, d)
    b_tri, b_t = ob.trace_brute(e, o, d)
    assert np.array_equal(w_tri, b_tri) and np.array_equal(w_t[w_tri >= 0], b_t[b_tri >= 0])
    f = cornell_frame(48, 48, 2, 5)
    _, hdr_a, info_a = ob.render_frame(e, f, 7)
    with ob.wide_tree(nodes, geo):
        _, hdr_b, info_b = ob.render_frame(e, f, 7)
    _, hdr_c, _ = ob.render_frame(e, f, 7)  # the setting ends with the block
    assert hdr_a.max() > 0
    assert np.array_equal(hdr_a.view(np.uint32), hdr_b.view(np.uint32))
    assert np.array_equal(hdr_a.view(np.uint32), hdr_c.view(np.uint32))
    assert info_a["rays"] == info_b["rays"]
    # the wide walk ran: one node visit per traced ray on the one-node tree
    assert info_b["node_visits"] == info_b["rays"] != info_a["node_visits"]
