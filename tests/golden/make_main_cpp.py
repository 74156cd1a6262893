"""Writes the generated C++ of tests/golden/main.ray with the reference's .ray
compiler (transpile, src/compiler/raymotion_compiler.py:8-54), plain and
preview variants, as fixtures for the caller-path test: the generated main()
must compile unchanged against include/raymotion and run through the HIP
backend. Run in the build container, where the reference tree exists:

    python tests/golden/make_main_cpp.py
"""
import importlib.util
from pathlib import Path

HERE = Path(__file__).resolve().parent
COMPILER = Path("/root/reference/src/compiler/raymotion_compiler.py")


def main():
    spec = importlib.util.spec_from_file_location("raymotion_compiler", COMPILER)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    source = (HERE / "main.ray").read_text()
    (HERE / "main_ray.cpp").write_text(mod.transpile(source, preview=False, filename="main.ray"))
    (HERE / "main_ray_preview.cpp").write_text(mod.transpile(source, preview=True, filename="main.ray"))


if __name__ == "__main__":
    main()
