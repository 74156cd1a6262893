"""GPU-built traversal tree (rm_set_scene_gpu_tree): a tree is correct when
every ray finds the closest hit, so the build is pinned by hit equality with
brute force (small scenes) and with the oracle's traversal of the host tree
(itself pinned to brute force), bit for bit in triangle id and distance; and
by images identical to renders with the host-built tree. Edge cases: one to
three triangles, thousands of coincident centroids (equal Morton codes), refit
after motion, and the rebuild after 31 refits."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import cornell, cornell_frame, frame, rel_err, tri
from raymotion_amd import Scene, scenes
from raymotion_amd.denoise import restir_frame
from raymotion_amd.rng import random_rays

pytestmark = pytest.mark.gpu


def _box(exported, pad=0.5):
    v = np.concatenate([exported.tris["v0"], exported.tris["v1"], exported.tris["v2"]]).reshape(-1, 3)
    return v.min(axis=0) - pad, v.max(axis=0) + pad


def _hits_equal(backend, exported, n_rays, seed, brute=False, any_hit=False):
    lo, hi = _box(exported)
    o, d = random_rays(n_rays, seed, lo=lo, hi=hi)
    g_tri, g_t, _, _ = backend.trace(o, d, None, any_hit)
    if brute:
        c_tri, c_t = ob.trace_brute(exported, o, d)
    else:
        c_tri, c_t, _, _, _ = ob.trace(exported, o, d, None, any_hit)
    if any_hit:
        return int(np.sum((g_tri >= 0) != (c_tri >= 0))), 0
    hit = c_tri >= 0
    t_bad = int(np.sum(g_t[hit].view(np.uint32) != c_t[hit].view(np.uint32)))
    return int(np.sum(g_tri != c_tri)), t_bad


def test_cornell_brute_force(backend):
    s = cornell()
    backend.set_scene(s, gpu_tree=True, rebuild=True)
    st = backend.stats()
    assert st.refit == 3 and st.bvh_nodes >= 1
    bad, t_bad = _hits_equal(backend, s.export(), 1 << 16, 11, brute=True)
    assert bad == 0 and t_bad == 0


@pytest.mark.parametrize("count", [1, 2, 3, 5])
def test_tiny_scenes(backend, count):
    s = Scene()
    rng = np.random.default_rng(count)
    for _ in range(count):
        a = rng.uniform(-1, 1, 3)
        tri(s, a, a + rng.uniform(0.2, 1, 3) * [1, 0, 0], a + rng.uniform(0.2, 1, 3) * [0, 1, 0], (0.5, 0.5, 0.5))
    s.build_bvh()
    backend.set_scene(s, gpu_tree=True)
    assert backend.stats().refit == 3
    bad, t_bad = _hits_equal(backend, s.export(), 1 << 14, count, brute=True)
    assert bad == 0 and t_bad == 0


def test_coincident_centroids(backend):
    """3000 copies of one triangle (equal Morton codes, equal distances: the
    smallest index wins) among 500 others."""
    s = Scene()
    rng = np.random.default_rng(5)
    for _ in range(3000):
        tri(s, (-1, -1, 2), (1, -1, 2), (0, 1, 2), (0.5, 0.5, 0.5))
    for _ in range(500):
        a = rng.uniform(-3, 3, 3)
        tri(s, a, a + [0.3, 0, 0], a + [0, 0.3, 0.1], (0.5, 0.5, 0.5))
    s.build_bvh()
    backend.set_scene(s, gpu_tree=True)
    st = backend.stats()
    print(f"coincident: {st.bvh_nodes} nodes, stack {st.stack_depth}")
    bad, t_bad = _hits_equal(backend, s.export(), 1 << 14, 3, brute=True)
    assert bad == 0 and t_bad == 0


def test_mixed_scene_oracle(backend):
    s = scenes.mixed_spheres(seed=0)
    s.build_bvh()
    e = s.export()
    backend.set_scene(e, gpu_tree=True)
    st = backend.stats()
    print(f"mixed: {st.bvh_nodes} nodes, stack {st.stack_depth}, build {st.scene_ms[1]:.2f} ms")
    for any_hit in (False, True):
        bad, t_bad = _hits_equal(backend, e, 1 << 17, 4, any_hit=any_hit)
        assert bad == 0 and t_bad == 0, any_hit


def _render(backend, scene, f, seed, gpu_tree):
    backend.set_scene(scene, gpu_tree=gpu_tree, rebuild=gpu_tree)
    rgb, hdr = backend.render(f, seed)
    st = backend.stats()
    return rgb, hdr, list(st.rays)


@pytest.mark.parametrize("which", ["restir", "megakernel"])
def test_images_equal_host_tree(backend, which):
    if which == "restir":
        s = scenes.mixed_spheres(seed=1, n_spheres=20, subdivisions=2)
        s.build_bvh()
        f = restir_frame(192, 108, 4, scenes.mixed_camera(3), denoise=1, temporal=0)
    else:
        s = cornell(rough=0.4)
        f = cornell_frame(96, 96, 8, 6)
    a = _render(backend, s, f, 31, False)
    b = _render(backend, s, f, 31, True)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert a[2] == b[2]  # same rays; node visits and triangle tests depend on the tree


def test_refit_and_rebuild_cadence(backend):
    """An animated height field: build, then refits that follow the vertices
    (hits against the oracle with a fresh host tree), and a rebuild after 31
    refits in a row."""
    n = 96
    s = scenes.heightfield(n, seed=2)
    anim = scenes.HeightFieldAnimation(n, seed=2)
    s.set_positions(anim.positions(0.0))
    backend.set_scene(s, gpu_tree=True, rebuild=True)
    assert backend.stats().refit == 3
    for k in range(1, 33):
        s.set_positions(anim.positions(k / 10.0))
        backend.set_scene(s, gpu_tree=True)
        st = backend.stats()
        assert st.refit == (3 if k == 32 else 1), (k, st.refit)
        if k in (1, 31, 32):
            s.build_bvh()
            bad, t_bad = _hits_equal(backend, s.export(), 1 << 15, k)
            assert bad == 0 and t_bad == 0, k


def test_full_heightfield_oracle(backend):
    """The 2,093,314-triangle animated height field: build time, tree shape and
    hit equality with the oracle for 2^17 rays."""
    s = scenes.heightfield(1024, seed=0)
    s.build_bvh()
    e = s.export()
    backend.set_profiling(True)
    try:
        backend.set_scene(e, gpu_tree=True, rebuild=True)
        st = backend.stats()
    finally:
        backend.set_profiling(False)
    print(f"heightfield: {st.bvh_nodes} nodes, stack {st.stack_depth}, build {st.scene_ms[1]:.2f} ms GPU")
    assert st.stack_depth <= 48  # the kernels with a 48-entry stack keep their occupancy
    bad, t_bad = _hits_equal(backend, e, 1 << 17, 9)
    assert bad == 0 and t_bad == 0
    # frames on this tree (over 2^19 nodes: the path engine's large-tree
    # instance) against the oracle's frames on the host tree
    assert st.bvh_nodes > 1 << 19
    cam = scenes.heightfield_camera(3)
    for name, f in (("restir", frame(96, 54, 4, 6, cam=cam, candidates=8, spatial=4, radius=16.0)),
                    ("megakernel", frame(96, 54, 4, 6, cam=cam))):
        g_rgb, g_hdr = backend.render(f, 4321)
        c_rgb, c_hdr, info = ob.render_frame(e, f, 4321)
        err = rel_err(g_hdr, c_hdr)
        print(f"heightfield {name} frame: path {info['path']}, max rel {err.max():.3g}, "
              f"exact {np.mean(g_hdr == c_hdr):.6f}")
        assert info["path"] == (1 if name == "restir" else 0)
        assert err.max() <= 1e-3


def _spread21(v):
    x = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    for shift, mask in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                        (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
        x = (x | (x << np.uint64(shift))) & np.uint64(mask)
    return x


def _morton_order(exported):
    """The leaf order the build must produce: triangles stably sorted by the
    63-bit Morton code of their centroid (v0 + v1 + v2, FP32) over the
    centroid bounds, 21 bits per axis interleaved x, y, z from the top."""
    t = exported.tris
    c = (t["v0"] + t["v1"]) + t["v2"]  # float32, the kernel's association
    lo, hi = c.min(axis=0), c.max(axis=0)
    q = []
    for a in range(3):
        span = hi[a] - lo[a]
        u = (c[:, a] - lo[a]) / span if span > 0 else np.zeros(len(c), np.float32)
        u = np.clip(u, np.float32(0), np.float32(1)).astype(np.float32)
        q.append(np.minimum((u * np.float32(2097152.0)).astype(np.uint32), np.uint32(2097151)))
    key = (_spread21(q[0]) << np.uint64(2)) | (_spread21(q[1]) << np.uint64(1)) | _spread21(q[2])
    return np.argsort(key, kind="stable"), key


@pytest.mark.parametrize("scene", ["mixed", "heightfield", "coincident"])
def test_leaf_order_is_stable_morton_sort(backend, scene):
    """The device radix sort of (Morton code, triangle) pairs: the built
    tree's leaf-order triangles equal a stable host sort of the same keys,
    ties (equal codes) in triangle order."""
    if scene == "mixed":
        s = scenes.mixed_spheres(seed=5, n_spheres=16, subdivisions=3)
    elif scene == "heightfield":
        s = scenes.heightfield(seed=2, n=256)
    else:
        s = Scene()
        for k in range(3000):  # many equal codes: a few distinct centroids
            a = np.array([k % 3, 0.0, 0.0])
            tri(s, a, a + [0.1, 0, 0], a + [0, 0.1, 0], (0.5, 0.5, 0.5))
    e = s.export()
    backend.set_scene(e, gpu_tree=True, rebuild=True)
    _, tris, _ = backend.read_tree()
    leaf_ids = tris[:, 9].view(np.int32)
    order, key = _morton_order(e)
    print(f"{scene}: {len(order)} triangles, {len(np.unique(key))} distinct codes", flush=True)
    assert np.array_equal(leaf_ids, order)
