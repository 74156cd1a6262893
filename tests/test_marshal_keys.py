"""The boundary's marshal (render_image_hip, rm_scene_upload) takes the
records' content keys while writing them and rewrites only records that
changed; tests/cpp/marshal_keys.cpp checks records and keys against a fresh
marshal and the backend's own hash of the records after each kind of scene
change. Host code only."""
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent / "cpp" / "_bin" / "marshal_keys"


def test_marshal_records_and_keys():
    if not EXE.exists():
        pytest.skip(f"{EXE} not built (make -C tests/cpp)")
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
