"""The product's host engine against the reference's own CPU engine
(scene.cpp, bvh.cpp, camera.cpp compiled from the reference sources into
oracle/_ref/libref.so): bit-identical triangle precompute, light CDF, BVH
nodes, triangle order, refit results and camera matrices. Also the checks of
the reference's tests/test_bvh.cpp."""
import numpy as np
import pytest

import oracle_bindings as ob
import raymotion_amd as rm
from raymotion_amd import scenes
from raymotion_amd.scene import material_rows

needs_ref = pytest.mark.skipif(ob.ref() is None, reason="reference engine not built (oracle/_ref)")


def _soup(n, seed, spread=10.0, emissive_every=7):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-spread, spread, (n, 1, 3))
    pos = centers + rng.normal(0.0, 0.5, (n, 3, 3))
    pos[::11, 2] = pos[::11, 1]  # some degenerate triangles (rejected by both)
    mat = material_rows(n, albedo=(0.5, 0.6, 0.7), rough=0.4)
    mat[::emissive_every, 3:6] = rng.uniform(0.5, 5.0, (len(mat[::emissive_every]), 3))
    mat[:, 9] = rng.choice([-1.0, 0.3, 1.0, 2.0], n)  # alpha clamping
    return pos, mat


def _both(pos, mat, normals=None, leaf=8):
    a = rm.Scene()
    ia = a.add_triangles(pos, mat, normals)
    b = ob.RefScene()
    ib = b.add_triangles(pos, mat, normals)
    assert np.array_equal(ia, ib)
    a.build_bvh(leaf)
    b.build_bvh(leaf)
    return a, b


def _assert_same(a, b):
    assert a.counts() == b.counts()
    for x, y in zip(a.precompute(), b.precompute()):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    for x, y in zip(a.nodes(), b.nodes()):
        assert np.array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("n,seed,leaf", [(1, 0, 8), (37, 1, 2), (1000, 2, 8), (20000, 3, 4), (20000, 4, 8)])
def test_random_soup_matches_reference(n, seed, leaf):
    pos, mat = _soup(n, seed)
    a, b = _both(pos, mat, leaf=leaf)
    _assert_same(a, b)


@needs_ref
def test_degenerate_and_coincident_geometry_matches_reference():
    # Zero-area node boxes and identical centroids exercise the sort fallbacks.
    pos = np.tile(np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float64), (50, 1, 1))
    pos[25:] += 1e-9
    mat = material_rows(50)
    a, b = _both(pos, mat, leaf=2)
    _assert_same(a, b)


@needs_ref
def test_mixed_scene_matches_reference():
    a = rm.Scene()
    b = ob.RefScene()
    for pos, mat, nrm in scenes.mixed_sphere_parts(seed=0, n_spheres=12):
        a.add_triangles(pos, mat, nrm)
        b.add_triangles(pos, mat, nrm)
    a.build_bvh()
    b.build_bvh()
    _assert_same(a, b)


@needs_ref
def test_refit_matches_reference():
    pos, mat = _soup(3000, 5)
    a, b = _both(pos, mat)
    moved = pos + np.random.default_rng(9).normal(0.0, 0.3, pos.shape)
    for sc in (a, b):
        sc.clear()
    a.add_triangles(moved, mat)
    b.add_triangles(moved, mat)
    a.refit_bvh()
    b.refit_bvh()
    _assert_same(a, b)


@needs_ref
@pytest.mark.parametrize("pos,rot,fov", [((0, 0, 0), (0, 0, 0), 60), ((1, 2, 3), (10, -35, 5), 40),
                                         ((0.5, 0.5, -1.4), (0, 0, 0), 40), ((-3, 1, 7), (80, 170, -45), 90)])
def test_camera_matches_reference(pos, rot, fov):
    c = scenes.camera(pos, rot, fov)
    want = ob.ref_camera(pos, rot, fov).astype(np.float32)
    got = np.array([c.x, c.y, c.z, c.focal, c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, c.m6, c.m7, c.m8], np.float32)
    assert np.array_equal(got, want)


def test_reference_test_bvh_cases():
    """tests/test_bvh.cpp:13-24 against the product engine."""
    def fill(s, z, n):
        s.clear()
        pos = np.array([[(3 * i - 1, -1, z), (3 * i + 1, -1, z), (3 * i, 1, z)] for i in range(n)], np.float64)
        nrm = np.tile(np.array([0.0, 0.0, -1.0]), (n, 3, 1))
        s.add_triangles(pos, material_rows(n, rough=0.5), normals=nrm)
    s = rm.Scene()
    fill(s, 3, 40)
    s.build_bvh(2)
    fill(s, 7, 40)
    s.refit_bvh()
    boxes, _, _ = s.nodes()
    assert boxes[0, 2] == 7 and boxes[0, 5] == 7, "refit bounds"
    r = rm.Scene()
    fill(r, 7, 40)
    r.build_bvh(2)
    rb, _, _ = r.nodes()
    assert boxes[0, 0] == rb[0, 0] and boxes[0, 3] == rb[0, 3], "refit differs from rebuild"
    fill(s, 9, 1)
    s.refit_bvh()
    boxes, _, bvh_tri = s.nodes()
    assert boxes[0, 2] == 9 and len(bvh_tri) == 1, "topology change"
    s.clear()
    s.refit_bvh()
    assert s.counts()[1] == 0, "empty scene"


def test_alpha_clamped_by_add_triangle():
    """tests/test_alpha.cpp:30-31: alpha -1 -> 0, 2 -> 1, NaN -> 1."""
    s = rm.Scene()
    pos = np.array([[(-10, -10, z), (0, 10, z), (10, -10, z)] for z in (1, 2, 3)], np.float64)
    mat = material_rows(3)
    mat[:, 9] = [-1.0, 2.0, np.nan]
    s.add_triangles(pos, mat)
    s.build_bvh()
    e = s.export()
    assert list(e.tris["alpha"]) == [0.0, 1.0, 1.0]


def test_export_preorder_escape_links():
    """metal.mm:129-144: preorder layout, right = index past the subtree."""
    s = scenes.cornell_box()
    s.build_bvh(2)
    e = s.export()
    n = len(e.nodes)
    assert e.nodes[0]["right"] == n

    def check(i):
        if e.nodes[i]["count"] > 0:
            assert e.nodes[i]["right"] == i + 1
            return i + 1
        end_left = check(i + 1)
        assert e.nodes[i + 1]["right"] == end_left
        end = check(end_left)
        assert e.nodes[i]["right"] == end
        return end
    assert check(0) == n
    leaves = e.nodes[e.nodes["count"] > 0]
    covered = np.concatenate([e.bvh_tri[o:o + c] for o, c in zip(leaves["offset"], leaves["count"])])
    assert np.array_equal(np.sort(covered), np.arange(len(e.tris)))
