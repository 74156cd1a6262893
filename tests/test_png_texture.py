"""PNG map_Kd textures (the reference loads them with stb_image as 8-bit RGB,
src/engine/obj.cpp:113-124) through the engine's own decoder
(raymotion_amd/csrc/engine/png.cpp). The PNGs are written here with Python's
zlib, so every case has a byte-exact expectation that does not come from the
decoder: each colour type and bit depth, the five row filters, Adam7
interlacing, stored / fixed-Huffman / dynamic-Huffman deflate blocks, image
data split over several IDAT chunks, and malformed files that must be
rejected; the decoder runs under AddressSanitizer and UBSan, also on
randomly corrupted files. Then an OBJ whose MTL names a PNG texture (quoted,
with a backslash) decodes it to the same sRGB-linearised texels as the
equivalent PPM."""
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
ENGINE = ROOT / "raymotion_amd" / "csrc" / "engine"

HARNESS = r'''
#include "png.h"
#include <cstdio>
#include <fstream>
#include <sstream>
int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        std::ifstream f(argv[i], std::ios::binary);
        std::ostringstream s;
        s << f.rdbuf();
        try {
            raymotion::Rgb8Image img = raymotion::decode_png(s.str());
            std::printf("%d %d ", img.width, img.height);
            for (unsigned char b : img.rgb) std::printf("%02x", b);
            std::printf("\n");
        } catch (const std::exception& e) {
            std::printf("error %s\n", e.what());
        }
    }
    return 0;
}
'''

OBJ_HARNESS = r'''
#include <raymotion/obj.hpp>
#include <cstdio>
using namespace raymotion;
int main(int, char** argv) {
    Scene a, b;
    if (load_obj_materials(a, argv[1], argv[2]) != 1 || load_obj_materials(b, argv[1], argv[3]) != 1) return 2;
    const Texture& ta = *a.tris[0].texture;
    const Texture& tb = *b.tris[0].texture;
    bool same = ta.width == tb.width && ta.height == tb.height && ta.pixels.size() == tb.pixels.size();
    for (size_t i = 0; same && i < ta.pixels.size(); ++i)
        same = ta.pixels[i].x == tb.pixels[i].x && ta.pixels[i].y == tb.pixels[i].y && ta.pixels[i].z == tb.pixels[i].z;
    std::printf("%d %d %s\n", ta.width, ta.height, same ? "same" : "differ");
    return same ? 0 : 1;
}
'''


@pytest.fixture(scope="module")
def decoder(tmp_path_factory):
    d = tmp_path_factory.mktemp("png")
    (d / "harness.cpp").write_text(HARNESS)
    exe = d / "decode"
    # under AddressSanitizer and UBSan: a read past the input or an overflow aborts
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I", str(ENGINE), str(d / "harness.cpp"),
                    str(ENGINE / "png.cpp"), "-o", str(exe)], check=True)

    def run(files):
        out = subprocess.run([str(exe), *map(str, files)], capture_output=True, text=True, check=True).stdout
        res = []
        for line in out.splitlines():
            if line.startswith("error"):
                res.append(line)
            else:
                w, h, hexdata = line.split()
                res.append(np.frombuffer(bytes.fromhex(hexdata), np.uint8).reshape(int(h), int(w), 3))
        return res
    return run


# ---- a small PNG writer (PNG specification: chunks, filters, Adam7)

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def pack_row(samples, depth):
    """samples: (w, channels) ints -> row bytes (big-endian 16-bit, MSB-first sub-byte)."""
    flat = samples.reshape(-1)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in flat)
    if depth == 8:
        return bytes(int(v) for v in flat)
    bits = "".join(format(int(v), f"0{depth}b") for v in flat)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else b if pb <= pc else c


def filter_row(raw, prev, ftype, bpp):
    out = bytearray([ftype])
    for i, x in enumerate(raw):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][ftype]
        out.append((x - pred) & 0xFF)
    return bytes(out)


def encode(samples, ctype, depth, interlace=0, filters=(0, 1, 2, 3, 4), palette=None, level=6, strategy=None,
           idat_split=1):
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw, k = b"", 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = [pack_row(r, depth) for r in sub]
        prev = bytes(len(rows[0]))
        for r in rows:
            raw += filter_row(r, prev, filters[k % len(filters)], bpp)
            prev, k = r, k + 1
    if strategy is None:
        z = zlib.compress(raw, level)
    else:
        co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
        z = co.compress(raw) + co.flush()
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if palette is not None:
        png += chunk(b"PLTE", bytes(palette.astype(np.uint8).reshape(-1)))
    png += chunk(b"tEXt", b"Comment\x00ancillary chunks are skipped")
    step = -(-len(z) // idat_split)
    for i in range(0, len(z), step):
        png += chunk(b"IDAT", z[i:i + step])
    return png + chunk(b"IEND", b"")


def expected_rgb(samples, ctype, depth, palette=None):
    s = samples.astype(np.int64)
    if depth == 16:
        s = s >> 8
    elif depth < 8 and ctype == 0:
        s = s * {1: 255, 2: 85, 4: 17}[depth]
    if ctype == 3:
        return palette[s[..., 0]].astype(np.uint8)
    if ctype in (0, 4):
        return np.repeat(s[..., :1], 3, axis=2).astype(np.uint8)
    return s[..., :3].astype(np.uint8)


def random_samples(rng, h, w, ctype, depth, ncolours=None):
    top = (ncolours or (1 << depth))
    return rng.integers(0, top, (h, w, CHANNELS[ctype]))


CASES = [  # (ctype, depth, interlace, w, h)
    (2, 8, 0, 13, 7), (2, 16, 0, 5, 4), (6, 8, 0, 9, 6), (6, 16, 0, 4, 3),
    (0, 1, 0, 11, 5), (0, 2, 0, 7, 3), (0, 4, 0, 5, 5), (0, 8, 0, 6, 6), (0, 16, 0, 3, 3),
    (4, 8, 0, 6, 4), (4, 16, 0, 3, 2),
    (3, 1, 0, 10, 3), (3, 2, 0, 9, 4), (3, 4, 0, 6, 3), (3, 8, 0, 8, 8),
    (2, 8, 1, 11, 9), (0, 2, 1, 10, 10), (6, 16, 1, 3, 5), (3, 4, 1, 1, 1), (2, 8, 1, 9, 1),
]


def test_png_colour_types_depths_filters_and_interlace(tmp_path, decoder):
    rng = np.random.default_rng(0)
    files, expect = [], []
    for i, (ctype, depth, interlace, w, h) in enumerate(CASES):
        palette = None
        if ctype == 3:
            n = min(1 << depth, 200)
            palette = rng.integers(0, 256, (n, 3))
            samples = random_samples(rng, h, w, ctype, depth, n)
        else:
            samples = random_samples(rng, h, w, ctype, depth)
        p = tmp_path / f"case{i}.png"
        p.write_bytes(encode(samples, ctype, depth, interlace, palette=palette, idat_split=1 + i % 3))
        files.append(p)
        expect.append(expected_rgb(samples, ctype, depth, palette))
    got = decoder(files)
    for case, g, e in zip(CASES, got, expect):
        assert not isinstance(g, str), (case, g)
        assert g.shape == e.shape and np.array_equal(g, e), case


@pytest.mark.parametrize("how", ["stored", "fixed", "dynamic", "one filter each"])
def test_png_deflate_block_kinds(tmp_path, decoder, how):
    rng = np.random.default_rng(1)
    # a smooth image compresses into long back-references
    y, x = np.mgrid[0:40, 0:48]
    samples = np.stack([(x * 5) % 256, (y * 3) % 256, ((x + y) * 2) % 256], axis=2)
    samples[::7, ::5] = rng.integers(0, 256, samples[::7, ::5].shape)
    kw = {"stored": {"level": 0}, "fixed": {"level": 9, "strategy": zlib.Z_FIXED},
          "dynamic": {"level": 9}, "one filter each": {"level": 6}}[how]
    files, expect = [], []
    for f in ((0,), (1,), (2,), (3,), (4,)) if how == "one filter each" else ((0, 1, 2, 3, 4),):
        p = tmp_path / f"img{len(files)}.png"
        p.write_bytes(encode(samples, 2, 8, filters=f, idat_split=4, **kw))
        files.append(p)
        expect.append(samples.astype(np.uint8))
    for g, e in zip(decoder(files), expect):
        assert not isinstance(g, str), g
        assert np.array_equal(g, e)


def test_png_rejects_malformed_files(tmp_path, decoder):
    rng = np.random.default_rng(2)
    samples = random_samples(rng, 4, 4, 2, 8)
    good = encode(samples, 2, 8)
    # 4x4 RGB8 rows whose filter byte is 5 (no such filter), in valid deflate data
    raw = b"".join(b"\x05" + bytes(12) for _ in range(4))
    ihdr = chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 4, 8, 2, 0, 0, 0))

    def png(idat):
        return b"\x89PNG\r\n\x1a\n" + ihdr + chunk(b"IDAT", idat) + chunk(b"IEND", b"")
    cases = {
        "truncated": good[:len(good) // 2],
        "signature": b"\x89PNX" + good[4:],
        "no palette": encode(random_samples(rng, 2, 2, 3, 8), 3, 8),
        "bad depth": good[:24] + bytes([3]) + good[25:],  # RGB at 3 bits per sample
        "bad filter": png(zlib.compress(raw)),
        "zlib header": png(b"\x78\x9d" + zlib.compress(raw)[2:]),  # header check fails
        "short data": png(zlib.compress(raw[:-5])),
    }
    files = []
    for name, data in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.png"
        p.write_bytes(data)
        files.append(p)
    for name, g in zip(cases, decoder(files)):
        assert isinstance(g, str) and g.startswith("error PNG:"), (name, g)


def test_png_corrupted_files_decode_or_fail_cleanly(tmp_path, decoder):
    """Byte flips, truncations and insertions in valid files of several kinds:
    each either decodes to an image of the header's size or is rejected, with
    no sanitizer report (the harness aborts on one)."""
    rng = np.random.default_rng(4)
    seeds = [encode(random_samples(rng, 9, 11, 2, 8), 2, 8, level=9),
             encode(random_samples(rng, 8, 8, 0, 2), 0, 2, interlace=1),
             encode(random_samples(rng, 6, 7, 3, 4, 16), 3, 4, palette=rng.integers(0, 256, (16, 3))),
             encode(random_samples(rng, 5, 5, 6, 16), 6, 16, strategy=zlib.Z_FIXED, level=9),
             encode(random_samples(rng, 12, 4, 2, 8), 2, 8, level=0)]
    files = []
    for i in range(400):
        b = bytearray(seeds[i % len(seeds)])
        for _ in range(1 + i % 4):
            at = int(rng.integers(8, len(b)))
            kind = int(rng.integers(3))
            if kind == 0:
                b[at] = int(rng.integers(256))
            elif kind == 1:
                del b[at:]
            else:
                b[at:at] = bytes(rng.integers(0, 256, int(rng.integers(1, 9))).astype(np.uint8))
            if len(b) <= 8:
                break
        p = tmp_path / f"fuzz{i}.png"
        p.write_bytes(bytes(b))
        files.append(p)
    got = decoder(files)
    assert len(got) == len(files)
    ok = sum(not isinstance(g, str) for g in got)
    print(f"{ok} of {len(files)} corrupted files still decode; the rest are rejected")
    assert all(isinstance(g, str) and g.startswith("error PNG:") or g.ndim == 3 for g in got)


def test_obj_png_texture_matches_ppm(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (5, 6, 3))
    (tmp_path / "tex").mkdir()
    (tmp_path / "tex" / "paint.png").write_bytes(encode(img, 2, 8))
    (tmp_path / "tex" / "paint.ppm").write_bytes(b"P6\n6 5\n255\n" + bytes(img.astype(np.uint8).reshape(-1)))
    (tmp_path / "model.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nusemtl m\nf 1/1 2/2 3/3\n")
    (tmp_path / "png.mtl").write_text('newmtl m\nKd 1 1 1\nmap_Kd "tex\\paint.png"\n')
    (tmp_path / "ppm.mtl").write_text("newmtl m\nKd 1 1 1\nmap_Kd tex/paint.ppm\n")
    (tmp_path / "h.cpp").write_text(OBJ_HARNESS)
    exe = tmp_path / "h"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "include"), str(tmp_path / "h.cpp"),
                    *(str(ENGINE / f) for f in ("obj.cpp", "png.cpp", "scene.cpp")), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(tmp_path / "model.obj"), str(tmp_path / "png.mtl"), str(tmp_path / "ppm.mtl")],
                       capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.split() == ["6", "5", "same"], r.stdout + r.stderr
