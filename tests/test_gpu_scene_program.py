"""The C++ caller path end to end on the GPU.

* A scene program written like the main() the .ray compiler generates
  (tests/cpp/scene_program.cpp: Objects from argv, push, render, finish)
  renders a frame through Renderer and render_image_hip and writes a PNG.
* The .ray compiler's own output for tests/golden/main.ray
  (src/compiler/raymotion_compiler.py:8-54; fixtures written by
  tests/golden/make_main_cpp.py), compiled unchanged against include/raymotion:
  the plain variant writes a PNG, the preview variant streams
  "RAY1 w h frame size" frames on descriptor 3 and takes one byte on stdin per
  further frame (include/raymotion/preview.hpp).

Every scene is an emissive quad (emission (1, 0, 0)) filling the view, so
every pixel is exactly (127, 0, 0), the reference's own known answer for
emission (tests/test_metal.cpp:47-52), with the megakernel path and with the
ReSTIR + denoiser path."""
import os
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_bin"
EXE = BIN / "scene_program"
QUAD_OBJ = "v -1 -1 0\nv -1 1 0\nv 1 1 0\nv 1 -1 0\nf 1 2 3\nf 1 3 4\n"

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


def _built(name):
    exe = BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (make -C tests/cpp)")
    return exe


@pytest.mark.parametrize("restir", ["0", "1"])
def test_generated_main_renders_png(tmp_path, restir):
    exe = _built("main_ray")
    (tmp_path / "quad.obj").write_text(QUAD_OBJ)
    out = tmp_path / "frame.png"
    r = subprocess.run([str(exe), str(out), "32", "24", "4", "30", "png", "1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120, env={**os.environ, "RAYMOTION_RESTIR": restir})
    assert r.returncode == 0, r.stderr
    img = _read_png(out)
    assert img.shape == (24, 32, 3)
    assert np.all(img[..., 0] == 127) and np.all(img[..., 1:] == 0), np.unique(img.reshape(-1, 3), axis=0)[:8]


def test_generated_preview_streams_frames(tmp_path):
    """One byte on stdin grants a second frame; end of input ends the program."""
    exe = _built("main_ray_preview")
    (tmp_path / "quad.obj").write_text(QUAD_OBJ)
    r = subprocess.run(["bash", "-c", f'"{exe}" unused 32 24 4 30 mp4 0 3>frames.bin'], cwd=tmp_path, input=b"g",
                       capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    data = (tmp_path / "frames.bin").read_bytes()
    frames = []
    while data:
        header, _, data = data.partition(b"\n")
        tag, w, h, idx, size = header.split()
        assert tag == b"RAY1" and (int(w), int(h), int(size)) == (32, 24, 32 * 24 * 3)
        frames.append((int(idx), np.frombuffer(data[:int(size)], np.uint8).reshape(24, 32, 3)))
        data = data[int(size):]
    assert [i for i, _ in frames] == [0, 1]
    for _, img in frames:
        assert np.all(img[..., 0] == 127) and np.all(img[..., 1:] == 0)
