"""The reference's own C++ tests (tests/test_bvh.cpp, tests/test_runtime.cpp),
compiled unchanged against this repository's kept API headers and engine by
oracle/Makefile into oracle/_ref/ where the reference sources are mounted.
test_runtime records the frames with its own Renderer; it runs with
RAYMOTION_GPU_BVH=0, the reference's host-built tree (with the default GPU
build the runtime asks the backend for the tree instead of calling
Renderer::build, which that test counts)."""
import os
import subprocess
from pathlib import Path

import pytest

REF_BIN = Path(__file__).resolve().parent.parent / "oracle" / "_ref"


def _run(name, tmp_path, **env):
    exe = REF_BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (the reference sources were absent at build time)")
    return subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                          env={**os.environ, **env})


def test_reference_test_bvh(tmp_path):
    r = _run("test_bvh", tmp_path)
    assert r.returncode == 0, r.stderr


def test_reference_test_runtime(tmp_path):
    r = _run("test_runtime", tmp_path, RAYMOTION_GPU_BVH="0")
    assert r.returncode == 0, r.stderr
    assert "runtime geometry passed" in r.stdout


def test_reference_test_metal_skips_without_gpu(tmp_path):
    # tests/test_metal.cpp:23-27 exits 77 (skip) when render_image_metal says
    # "No Metal GPU found"; on a host without a GPU the backend says so.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present (tests/test_gpu_reference_cpp.py runs the binary there)")
    r = _run("test_metal", tmp_path)
    assert r.returncode == 77, (r.returncode, r.stderr)
    assert "No Metal GPU found" in r.stderr
