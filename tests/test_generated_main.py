"""The .ray compiler's output for tests/golden/main.ray
(src/compiler/raymotion_compiler.py:8-54) compiles unchanged against
include/raymotion and the HIP library (tests/cpp/Makefile), and its argument
check runs before anything touches the GPU. The frames it renders are checked
on the GPU in tests/test_gpu_scene_program.py."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_bin"
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_generated_sources_keep_the_program():
    # the compiler turns `void quad = object.init(` into `auto quad = ...`,
    # leaves the same text inside a string alone and hoists the #include
    plain = (GOLDEN / "main_ray.cpp").read_text()
    preview = (GOLDEN / "main_ray_preview.cpp").read_text()
    assert "auto quad = object.init(\"quad.obj\");" in plain
    assert "\"strings keep void x = object.init( untouched\"" in plain
    assert plain.index("#include <vector>") < plain.index("int main(")
    assert "#include <raymotion/preview.hpp>" in preview and "PreviewStream stream;" in preview


@pytest.mark.parametrize("name", ["main_ray", "main_ray_preview"])
def test_generated_program_checks_its_arguments(name):
    exe = BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (make -C tests/cpp)")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "raymotion: run this program with raymotion export" in r.stderr
