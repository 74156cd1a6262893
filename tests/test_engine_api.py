"""Host-side API checks that need no GPU: the C-ABI exports every declared
symbol, render switches map to frame settings as metal.mm:146-170 does, the
runtime keeps the reference's scene-program semantics (C++ test with a
recording renderer), and a scene program compiles against the kept headers."""
import ctypes
import os
import re
import subprocess
from pathlib import Path

import pytest

import raymotion_amd as rm
from raymotion_amd import _capi

ROOT = Path(__file__).resolve().parent.parent
ENGINE = ROOT / "raymotion_amd" / "csrc" / "engine"


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(rm_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    declared = _declared_symbols()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared == set(_capi.EXPORTED), declared ^ set(_capi.EXPORTED)


def test_no_gpu_reports_clean_error():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(rm.BackendError) as e:
        rm.Backend(0)
    assert e.value.code == _capi.RM_ERR_NO_DEVICE
    assert "No HIP GPU found" in str(e.value)


ENV = ("RAYMOTION_RESTIR", "RAYMOTION_DENOISE", "RAYMOTION_RESTIR_TEMPORAL", "RAYMOTION_RESTIR_CANDIDATES",
       "RAYMOTION_RESTIR_SPATIAL", "UOW2_NOADAPT", "UOW2_NOMIS")


@pytest.fixture
def clean_env(monkeypatch):
    for k in ENV:
        monkeypatch.delenv(k, raising=False)
    return monkeypatch


def test_frame_defaults(clean_env):
    f = rm.default_frame(640, 360, 16, 5)
    assert (f.restir_candidates, f.restir_spatial, f.restir_mcap, f.restir_radius) == (8, 4, 128, 16.0)
    assert (f.adapt_min, f.adapt_step, f.denoise, f.temporal, f.nomis) == (16, 8, 1, 1, 0)
    assert abs(f.clamp_value - 16) < 1e-6 and f.denoise_iterations == 2


def test_frame_env_switches(clean_env):
    clean_env.setenv("RAYMOTION_RESTIR", "0")
    f = rm.default_frame(64, 64, 4, 5)
    assert (f.restir_candidates, f.restir_spatial) == (1, 0)  # legacy NEE, no denoiser
    clean_env.setenv("RAYMOTION_RESTIR", "1")
    clean_env.setenv("RAYMOTION_DENOISE", "0")
    clean_env.setenv("RAYMOTION_RESTIR_TEMPORAL", "0")
    clean_env.setenv("UOW2_NOADAPT", "1")
    clean_env.setenv("UOW2_NOMIS", "1")
    f = rm.default_frame(64, 64, 4, 5)
    assert (f.denoise, f.temporal, f.adapt_min, f.nomis) == (0, 0, 0, 1)
    clean_env.setenv("RAYMOTION_RESTIR_CANDIDATES", "100")
    clean_env.setenv("RAYMOTION_RESTIR_SPATIAL", "-3")
    f = rm.default_frame(64, 64, 4, 5)
    assert (f.restir_candidates, f.restir_spatial) == (100, 0)  # kernels clamp to 32 / 8


def _gxx(*args, out):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-I", str(ROOT / "include"), *map(str, args),
                    "-o", str(out)], check=True)


def test_runtime_semantics_cpp(tmp_path):
    exe = tmp_path / "test_runtime"
    _gxx(ROOT / "tests" / "cpp" / "test_runtime.cpp",
         *(ENGINE / f for f in ("scene.cpp", "bvh.cpp", "camera.cpp", "obj.cpp", "png.cpp")), out=exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "runtime checks passed" in r.stdout


def test_scene_program_compiles_against_kept_api(tmp_path):
    exe = tmp_path / "scene_program"
    _gxx(ROOT / "tests" / "cpp" / "scene_program.cpp", "-L", ROOT / "raymotion_amd", "-lrm_hip",
         f"-Wl,-rpath,{ROOT / 'raymotion_amd'}", out=exe)
    assert exe.exists()


def test_obj_loader(tmp_path):
    """OBJ + MTL + PPM texture through the engine (src/engine/obj.cpp behaviour:
    fan order, z flip, material switching, sRGB decode, UV orientation)."""
    (tmp_path / "m").mkdir()
    (tmp_path / "model.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nusemtl painted\n"
                                        "f 1/1 2/2 3/3\nusemtl light\nf -3 -2 -1\n")
    (tmp_path / "m" / "model.mtl").write_text("newmtl painted\nKd .5 .8 1\nPr .2\nPm .7\nmap_Kd color image.ppm\n"
                                              "newmtl light\nKe 2 1 0\nNi 1.5\n")
    (tmp_path / "m" / "color image.ppm").write_bytes(b"P6\n2 2\n255\n" + bytes([255, 0, 0, 0, 255, 0, 0, 0, 255,
                                                                                  128, 128, 128]))
    src = tmp_path / "obj_test.cpp"
    src.write_text(r'''
#include <raymotion/obj.hpp>
#include <cstdio>
#include <stdexcept>
using namespace raymotion;
int main(int, char** argv) {
    Scene s;
    int n = load_obj_materials(s, argv[1], argv[2]);
    const auto& t = s.tris[0];
    bool ok = n == 2 && s.light_tri.size() == 1 && t.ar == .5 && t.rough == .2 && t.metallic == .7 && t.texture &&
              !s.tris[1].texture && s.tris[1].ior == 1.5 && t.v1.x == 0 && t.v1.y == 1 && t.v2.x == 1 &&
              t.texture->sample(.25, .25).z == 1 && t.texture->sample(.25, .75).x == 1 &&
              t.texture->sample(-.75, 1.25).z == 1;
    double gray = t.texture->sample(.75, .25).x;
    ok = ok && gray > .215 && gray < .217;
    std::printf(ok ? "ok\n" : "bad\n");
    return ok ? 0 : 1;
}''')
    exe = tmp_path / "obj_test"
    _gxx(src, ENGINE / "obj.cpp", ENGINE / "png.cpp", ENGINE / "scene.cpp", out=exe)
    r = subprocess.run([str(exe), str(tmp_path / "model.obj"), str(tmp_path / "m" / "model.mtl")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_set_positions_matches_fresh_scene():
    """rm_scene_set_positions (parallel precompute, lights rebuilt in order)
    leaves the same records, light list and CDF as adding the moved triangles
    to an empty scene."""
    import numpy as np
    from raymotion_amd import scenes
    from raymotion_amd.scene import material_rows
    n, seed = 48, 3
    a = scenes.heightfield(n, seed)
    pos = scenes.HeightFieldAnimation(n, seed).positions(0.37)
    rng = np.random.default_rng(1)
    nrm = rng.normal(size=pos.shape)
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    a.set_positions(pos, nrm)
    rough, light_tri, light_mat = scenes._heightfield_extras(n, seed)
    g = len(pos) - len(light_tri)
    mat = material_rows(g, albedo=(0.7, 0.7, 0.7), metallic=1.0)
    mat[:, 8] = rough
    b = rm.Scene()
    b.add_triangles(pos[:g], mat, normals=nrm[:g])
    b.add_triangles(pos[g:], light_mat, normals=nrm[g:])
    ea, eb = a.export(), b.export()
    assert ea.tris.tobytes() == eb.tris.tobytes()
    assert np.array_equal(ea.light_tri, eb.light_tri)
    assert ea.light_cdf.tobytes() == eb.light_cdf.tobytes() and ea.light_total == eb.light_total
    with pytest.raises(rm.BackendError):  # a degenerate triangle rejects the whole update
        bad = pos.copy()
        bad[5, 1] = bad[5, 0]
        a.set_positions(bad)
    assert a.export().tris.tobytes() == ea.tris.tobytes()
