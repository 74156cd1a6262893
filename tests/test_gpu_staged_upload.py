"""GPU-tree uploads of an animated Scene (rm_scene_upload_gpu_tree) copy the
device-format geometry to staging buffers chunk by chunk while the scene is
marshalled, then take the staged buffers; uploads of the same records through
the public rm_set_scene_gpu_tree pack and copy the geometry after comparing
keys. Both must leave the same scene resident: frame by frame the images are
equal bit for bit, through frames with moved vertices (staged), unchanged
frames (nothing staged), the rebuild after 31 refits, and a rebuild for a
larger triangle count (staged geometry of the new count)."""
import numpy as np
import pytest

import raymotion_amd as rm
from helpers import frame
from raymotion_amd import scenes

pytestmark = pytest.mark.gpu


def test_staged_geometry_equals_packed(backend):
    n = 48
    s = scenes.heightfield(n, seed=3)
    anim = scenes.HeightFieldAnimation(n, seed=3)
    other = rm.Backend(0)
    try:
        f = frame(64, 48, 2, 4, cam=scenes.heightfield_camera(2), candidates=4, spatial=2, radius=8.0)
        refits = []
        for k in range(36):
            if k not in (5, 6, 20):  # these frames leave the vertices where they were
                s.set_positions(anim.positions(k / 10.0))
            backend.set_scene(s, gpu_tree=True)           # marshal with staging
            other.set_scene(s.export(), gpu_tree=True)    # records through the public entry point
            ra, ha = backend.render(f, 100 + k)
            rb, hb = other.render(f, 100 + k)
            refits.append(backend.stats().refit)
            assert backend.stats().refit == other.stats().refit, k
            assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32)), k
            assert np.array_equal(ra, rb), k
        print("refit codes:", refits, flush=True)
        assert 3 in refits[1:] and 1 in refits and 0 in refits  # a rebuild, refits and unchanged frames
        # More triangles (the staged geometry then feeds a rebuild), then a
        # frame that moves the original vertices again.
        rng = np.random.default_rng(4)
        a = rng.uniform(-6, 6, (40, 3)) + [0.0, 2.0, 0.0]
        extra = np.stack([a, a + [0.4, 0.0, 0.0], a + [0.0, 0.3, 0.4]], axis=1)
        s.add_triangles(extra, scenes.material_rows(len(extra), albedo=(0.6, 0.5, 0.4)))
        for k in (40, 41):
            if k == 41:
                s.set_positions(np.concatenate([anim.positions(4.1), extra]))
            backend.set_scene(s, gpu_tree=True)
            other.set_scene(s.export(), gpu_tree=True)
            assert backend.stats().refit == other.stats().refit == (3 if k == 40 else 1), k
            ra, ha = backend.render(f, 100 + k)
            rb, hb = other.render(f, 100 + k)
            assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32)), k
    finally:
        other.close()
