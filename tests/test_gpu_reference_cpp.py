"""The reference's own GPU tests (tests/test_metal.cpp, tests/test_alpha.cpp),
compiled unchanged against this repository's kept API and linked with the HIP
backend (oracle/Makefile builds them into oracle/_ref/ where the reference
sources are mounted): every known answer they check (empty scene black, the
emissive quad exactly (127, 0, 0) with adaptive sampling after a re-upload,
alpha coverage 0 / 85 +-3 / 127, the two-layer closed forms, 64 transparent
layers, determinism, 34 frames of refits and odd dispatch tails) through
render_image_metal."""
import subprocess
from pathlib import Path

import pytest

REF_BIN = Path(__file__).resolve().parent.parent / "oracle" / "_ref"

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["test_metal", "test_alpha"])
def test_reference_gpu_test(tmp_path, name):
    exe = REF_BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (the reference sources were absent at build time)")
    print(f"running {name}", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, f"{name} exited {r.returncode}: {r.stderr}"
