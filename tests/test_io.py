"""File formats of the reference's example pipeline (host code, no GPU):
the .c_unary decoder and -log (load_unary), the BMP reader (load_image),
the MSRC-coloured PNG / PPM writer (save_map) and the superpixel terms of
addSuperPixel for a given segmentation.

Pinning: the unaries decoded from the reference's own data/img.c_unary
must equal, bit for bit, the compiled reference's ProbImage decode (SHA-256
in data_img.json); the superpixel terms must equal the compiled
reference's exp_of_superpixels_ for its own mean-shift segmentation of
data/img.bmp within 1e-6 relative (Eigen evaluates that exp with its own
vectorised polynomial, one ulp from the C library's); the BMP reader and the
PNG writer are checked by round trips through independent Python code."""

import hashlib
import json
import os
import struct
import zlib

import numpy as np
import pytest

from conftest import GOLDEN

DATA = os.path.join(GOLDEN, "data")


def _fixture():
    with open(os.path.join(GOLDEN, "data_img.json")) as f:
        return json.load(f)


def test_load_unary_matches_reference_decode():
    import densecrf_ho_amd as D
    ref = _fixture()
    U, W, H = D.load_unary(os.path.join(DATA, "img.c_unary"))
    assert (W, H, U.shape[1]) == (ref["W"], ref["H"], ref["L"])
    assert hashlib.sha256(U.tobytes()).hexdigest() == ref["unary_sha256"]
    assert [float(U[i, k]) for i, k in [(0, 0), (1234, 5), (68159, 20)]] == ref["unary_samples"]


def test_load_unary_rejects_truncated_file(tmp_path):
    import densecrf_ho_amd as D
    with open(os.path.join(DATA, "img.c_unary"), "rb") as f:
        b = f.read()
    p = tmp_path / "short.c_unary"
    p.write_bytes(b[: len(b) // 2])
    with pytest.raises(D.EngineError, match="truncated"):
        D.load_unary(str(p))


def test_load_image_matches_fixture():
    import densecrf_ho_amd as D
    ref = _fixture()
    im, W, H = D.load_image(os.path.join(DATA, "img.bmp"))
    assert (W, H) == (ref["W"], ref["H"])
    assert hashlib.sha256(im.tobytes()).hexdigest() == ref["image_sha256"]


def _write_bmp(path, rgb, bpp, top_down):
    H, W, _ = rgb.shape
    stride = (W * bpp // 8 + 3) // 4 * 4
    rows = []
    for y in (range(H) if top_down else range(H - 1, -1, -1)):
        px = rgb[y][:, ::-1]  # B, G, R
        if bpp == 32:
            px = np.concatenate([px, np.full((W, 1), 255, np.uint8)], 1)
        r = px.astype(np.uint8).tobytes()
        rows.append(r + b"\0" * (stride - len(r)))
    data = b"".join(rows)
    hdr = struct.pack("<2sIHHI", b"BM", 54 + len(data), 0, 0, 54)
    info = struct.pack("<IiiHHIIiiII", 40, W, -H if top_down else H, 1, bpp, 0, len(data), 2835, 2835, 0, 0)
    path.write_bytes(hdr + info + data)


@pytest.mark.parametrize("bpp,top_down", [(24, False), (24, True), (32, False), (32, True)])
def test_load_image_bmp_variants(tmp_path, bpp, top_down):
    import densecrf_ho_amd as D
    rgb = np.random.default_rng(bpp + top_down).integers(0, 256, (7, 13, 3), dtype=np.uint8)  # odd width: row padding
    p = tmp_path / "x.bmp"
    _write_bmp(p, rgb, bpp, top_down)
    im, W, H = D.load_image(str(p))
    assert (W, H) == (13, 7)
    np.testing.assert_array_equal(im.reshape(H, W, 3), rgb)


def _read_png(path):
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(b):
        n, = struct.unpack(">I", b[pos:pos + 4])
        typ, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", b[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + data) & 0xffffffff
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", data)
        elif typ == b"IDAT":
            idat += data
        pos += 12 + n
    W, H, depth, ctype = ihdr[:4]
    assert depth == 8 and ctype == 2
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(H, 1 + 3 * W)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(H, W, 3)


MSRC = np.array([[128, 0, 0], [0, 128, 0], [128, 128, 0], [0, 0, 128], [0, 128, 128], [128, 128, 128], [192, 0, 0],
                 [64, 128, 0], [192, 128, 0], [64, 0, 128], [192, 0, 128], [64, 128, 128], [192, 128, 128],
                 [0, 64, 0], [128, 64, 0], [0, 192, 0], [128, 64, 128], [0, 192, 128], [128, 192, 128], [64, 64, 0],
                 [192, 64, 0], [0, 0, 0]], np.uint8)  # the MSRC dataset's colour code


@pytest.mark.parametrize("W,H", [(37, 11), (300, 250)])  # the second spans several stored deflate blocks
def test_save_map_png_and_ppm(tmp_path, W, H):
    import densecrf_ho_amd as D
    Q = np.random.default_rng(W).random((W * H, 21)).astype(np.float32)
    D.save_map(Q, W, H, str(tmp_path / "m.png"))
    np.testing.assert_array_equal(_read_png(tmp_path / "m.png"), MSRC[Q.argmax(1)].reshape(H, W, 3))
    D.save_map(Q, W, H, str(tmp_path / "m.ppm"))
    b = (tmp_path / "m.ppm").read_bytes()
    hdr = f"P6\n{W} {H}\n255\n".encode()
    assert b.startswith(hdr)
    np.testing.assert_array_equal(np.frombuffer(b[len(hdr):], np.uint8).reshape(H, W, 3),
                                  MSRC[Q.argmax(1)].reshape(H, W, 3))


def test_superpixel_terms_match_reference():
    """exp_of_superpixels_ / constants_ of the reference's addSuperPixel for its
    mean-shift segmentation of data/img.bmp (kept as a region map)"""
    import densecrf_ho_amd as D
    g = np.load(os.path.join(GOLDEN, "data_img_labels.npz"))
    im, _, _ = D.load_image(os.path.join(DATA, "img.bmp"))
    e, c = D.superpixel_terms(im, g["regions"], 50.0, 1000.0)
    np.testing.assert_allclose(e, g["exp_of_superpixels"], rtol=1e-6)
    np.testing.assert_array_equal(c, g["constants"])
