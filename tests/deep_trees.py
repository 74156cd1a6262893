"""Hand-made binary trees in the reference's upload format (preorder nodes,
`right` = escape index, leaf ranges into bvh_tri; src/engine/metal.mm:129-144)
whose 4-wide collapse needs a deep traversal stack or is very deep.

Triangles are half squares in planes x = c (vertices (c,-1,-1), (c,1,-1),
(c,-1,1), normal +x), so a ray along -x at (y, z) with y + z > 0 passes every
one of them.

* comb(depth): spine node s_k (k = 0 .. depth-1) has the side subtree q_k
  (an inner node over two one-triangle leaves at x = k + 0.25, k + 0.75) and
  s_{k+1}; s_depth is a leaf with one emissive triangle at x = depth + 0.5.
  Collapsing opens the spine (the larger box) twice per wide node, so each
  wide node holds q_k, q_{k+1}, q_{k+2} and s_{k+3}: three stacked inner
  children per level, a stack bound of about `depth` entries. A ray along -x
  from beyond the spine's end enters s_{k+3} first and pushes the three q's
  at every level.
* caterpillar(depth): spine node s_k has a one-triangle leaf at x = k + 0.5
  and s_{k+1}; the wide tree is about depth / 3 levels deep with no stack
  (one inner child per wide node).
"""
from __future__ import annotations

import numpy as np

import raymotion_amd as rm
from raymotion_amd._capi import NODE_DTYPE
from raymotion_amd.scene import Exported, material_rows

PAD = 1e-3  # boxes are padded so no box test is tighter than its triangles


def _half_square(c: float) -> np.ndarray:
    return np.array([[c, -1.0, -1.0], [c, 1.0, -1.0], [c, -1.0, 1.0]])


def _scene(planes: list[float], emissive_last: bool):
    sc = rm.Scene()
    pos = np.array([_half_square(c) for c in planes])
    mat = material_rows(len(planes), albedo=(0.7, 0.6, 0.5))
    if emissive_last:
        mat[-1, 0:3] = (0.5, 0.5, 0.5)
        mat[-1, 3:6] = (4.0, 3.0, 2.0)
    n = np.array([[[1.0, 0.0, 0.0]] * 3] * len(planes))
    idx = sc.add_triangles(pos, mat, normals=n)
    assert (idx == np.arange(len(planes))).all()
    sc.build_bvh()  # the engine's own tree; replaced below
    return sc, sc.export()


def _node(lo_x, hi_x, right, offset=0, count=0):
    return ((lo_x - PAD, -1.0 - PAD, -1.0 - PAD), (hi_x + PAD, 1.0 + PAD, 1.0 + PAD), right, offset, count)


def _finish(e: Exported, rows) -> Exported:
    nodes = np.zeros(len(rows), NODE_DTYPE)
    for i, (mn, mx, right, off, cnt) in enumerate(rows):
        nodes[i] = (mn, mx, right, off, cnt)
    return Exported(e.tris, nodes, np.arange(len(e.tris), dtype=np.int32), e.light_tri, e.light_cdf,
                    e.light_total, e.texels)


def comb(depth: int) -> Exported:
    planes = []
    for k in range(depth):
        planes += [k + 0.25, k + 0.75]
    planes.append(depth + 0.5)
    _, e = _scene(planes, emissive_last=True)
    total = 4 * depth + 1
    rows = []
    for k in range(depth):
        i = len(rows)
        assert i == 4 * k
        rows.append(_node(k + 0.25, depth + 0.5, total))             # s_k
        rows.append(_node(k + 0.25, k + 0.75, i + 4))                # q_k
        rows.append(_node(k + 0.25, k + 0.25, i + 3, 2 * k, 1))      # leaf a_k
        rows.append(_node(k + 0.75, k + 0.75, i + 4, 2 * k + 1, 1))  # leaf b_k
    rows.append(_node(depth + 0.5, depth + 0.5, total, 2 * depth, 1))  # s_depth
    return _finish(e, rows)


def caterpillar(depth: int) -> Exported:
    planes = [k + 0.5 for k in range(depth + 1)]
    _, e = _scene(planes, emissive_last=True)
    total = 2 * depth + 1
    rows = []
    for k in range(depth):
        i = len(rows)
        rows.append(_node(k + 0.5, depth + 0.5, total))       # s_k
        rows.append(_node(k + 0.5, k + 0.5, i + 2, k, 1))     # leaf
    rows.append(_node(depth + 0.5, depth + 0.5, total, depth, 1))
    return _finish(e, rows)


def rays(depth: int, n: int, seed: int):
    """Rays for the comb / caterpillar of `depth`: half along -x from beyond
    the spine's end (y, z in [-0.95, 0.95]^2, so about half of them pass
    every triangle), half from random points in random directions."""
    rng = np.random.default_rng(seed)
    h = n // 2
    o1 = np.column_stack([np.full(h, depth + 2.0), rng.uniform(-0.95, 0.95, (h, 2))])
    d1 = np.column_stack([-np.ones(h), rng.uniform(-1e-3, 1e-3, (h, 2))])
    o2 = np.column_stack([rng.uniform(-1.0, depth + 2.0, n - h), rng.uniform(-1.5, 1.5, (n - h, 2))])
    d2 = rng.normal(size=(n - h, 3))
    d = np.vstack([d1, d2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.vstack([o1, o2]).astype(np.float32), d.astype(np.float32)


def camera(depth: int, width: int):
    """Pinhole at x = depth + 3 looking down -x (image x along +z, y up)."""
    from helpers import default_camera
    c = default_camera(width)
    c.x = depth + 3.0
    c.m0, c.m1, c.m2 = 0.0, 0.0, -1.0
    c.m3, c.m4, c.m5 = 0.0, 1.0, 0.0
    c.m6, c.m7, c.m8 = 1.0, 0.0, 0.0
    return c
