"""Host-side pieces of the adjoint and denoiser wrappers (no GPU): the
parameter layout agrees with the library's, initialisation leaves padding at
zero, and DiffScene maps parameters onto the exported scene (materials, and a
light CDF that reproduces the engine's at the initial parameters)."""
import numpy as np

from raymotion_amd import denoise, scenes
from raymotion_amd.diff import DiffScene


def test_kpn_layout_matches_library():
    assert denoise.param_count() == sum(L.cout * L.kp + L.cout for L in denoise.layout())
    p = denoise.init_params(3)
    for i, L in enumerate(denoise.layout()):
        w = p[L.w_off:L.w_off + L.cout * L.kp].reshape(L.cout, L.kp)
        assert np.all(w[:, 9 * L.cin:] == 0)          # K padding
        if i == len(denoise.layout()) - 1:
            assert np.all(w[denoise.TAPS:] == 0)      # unused logits
        assert np.all(p[L.b_off:L.b_off + L.cout] == 0)


def test_diffscene_roundtrip():
    s, groups = scenes.inverse_scene(seed=0, n_spheres=20, subdivisions=1)
    e = s.export()
    ds = DiffScene(e, groups)
    x = ds.exported(ds.initial)
    for f in ("ar", "ag", "ab", "rough", "er", "eg", "eb"):
        np.testing.assert_array_equal(x.tris[f], e.tris[f])
    np.testing.assert_allclose(x.light_cdf, e.light_cdf, rtol=1e-6)
    assert abs(x.light_total - e.light_total) <= 1e-6 * e.light_total
    p = ds.initial.copy()
    p[14, 4] = 2.0                       # double one light group's emission
    y = ds.exported(p)
    sel = groups == 14
    np.testing.assert_allclose(y.tris["er"][sel], 2 * e.tris["er"][sel], rtol=1e-6)
    assert y.light_total > e.light_total
    q = ds.clamp(np.full_like(p, 5.0))
    assert np.all(q[:, :4] == 1.0) and np.all(q[:, 4] == 5.0)
