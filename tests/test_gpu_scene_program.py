"""The C++ caller path end to end on the GPU: a scene program written like the
main() the .ray compiler generates (tests/cpp/scene_program.cpp: Objects from
argv, push, render, finish) renders a frame through Renderer and
render_image_hip and writes a PNG. Its scene is an emissive quad (emission
(1, 0, 0)) filling the view, so every pixel is exactly (127, 0, 0), the
reference's own known answer for emission (tests/test_metal.cpp:47-52), with
the megakernel path and with the ReSTIR + denoiser path."""
import os
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

EXE = Path(__file__).resolve().parent / "cpp" / "_bin" / "scene_program"

pytestmark = pytest.mark.gpu


def _read_png(path):
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)  # filter type 0 on every row
    return raw[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("restir", ["0", "1"])
def test_scene_program_renders_png_through_hip(tmp_path, restir):
    if not EXE.exists():
        pytest.skip(f"{EXE} not built (make -C tests/cpp)")
    out = tmp_path / "frame.png"
    r = subprocess.run([str(EXE), str(out), "32", "24", "4", "30", "png", "1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120, env={**os.environ, "RAYMOTION_RESTIR": restir})
    assert r.returncode == 0, r.stderr
    img = _read_png(out)
    assert img.shape == (24, 32, 3)
    assert np.all(img[..., 0] == 127) and np.all(img[..., 1:] == 0), np.unique(img.reshape(-1, 3), axis=0)[:8]
