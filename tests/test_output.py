"""The output stage after the accumulation buffer: the oracle's 8-bit
conversion and auto exposure against closed forms, and the library's PNG
writer read back by an independent decoder (zlib + the PNG chunk rules).
CPU only; the GPU versions are compared with the oracle in
test_gpu_output.py.

Reference: computeAutoExposure / logLuminanceSumKernel
(src/path_tracer.cu:3101-3124), main's output conversion (:4366-4380),
saveImageRgb8Png (src/image_io.cpp:62-65).
"""
import os
import struct
import sys
import zlib

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle  # noqa: E402


def test_auto_exposure_closed_form():
    # constant radiance L: exposure = 0.18 / (1e-4 + L)
    for L, spp in ((0.5, 4), (2.0, 1), (0.0, 3)):
        accum = np.full((16, 24, 3), L * spp, np.float32)
        got = pyoracle.auto_exposure(accum, spp)
        assert got == pytest.approx(0.18 / (1e-4 + L), rel=2e-6)
    # two halves: geometric mean of the two luminances (+1e-4 each)
    accum = np.zeros((8, 8, 3), np.float32)
    accum[:4] = 0.25
    accum[4:] = 4.0
    want = 0.18 / np.exp(0.5 * (np.log(1e-4 + 0.25) + np.log(1e-4 + 4.0)))
    assert pyoracle.auto_exposure(accum, 1) == pytest.approx(want, rel=2e-6)
    # non-finite and negative luminance count as 0
    accum[0, 0] = np.nan
    accum[0, 1] = -3.0
    accum[0, 2] = np.inf
    zero = np.log(1e-4)
    terms = np.where(np.arange(64).reshape(8, 8) < 32, np.log(1e-4 + 0.25), np.log(1e-4 + 4.0))
    terms[0, :3] = zero
    assert pyoracle.auto_exposure(accum, 1) == pytest.approx(0.18 / np.exp(terms.mean()), rel=2e-6)


def test_rgb8_conversion_closed_values():
    # no tone map, exposure 1: linear -> sRGB OETF -> (uint8)(255.99 x)
    lin = np.array([0.0, 0.001, 0.0031308, 0.2140411, 0.5, 1.0, 2.0, np.nan, -1.0, np.inf], np.float32)
    accum = np.repeat(lin[:, None], 3, 1).reshape(1, -1, 3)
    got = pyoracle.resolve_rgb8(accum, 1)[0, :, 0]

    def oetf(x):
        x = max(x, 0.0)
        return 12.92 * x if x <= 0.0031308 else 1.055 * x ** (1 / 2.4) - 0.055

    for x, g in zip(lin, got):
        if not np.isfinite(x):
            want = 0  # NaN and +inf reach the shield as non-finite: 0
        else:
            want = int(255.99 * min(max(oetf(float(x)), 0.0), 1.0))
        assert g == want, (x, g, want)
    assert got[3] == 127  # sRGB(0.2140411) = 0.5 -> 127.995 -> 127


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        chunks.append((kind, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for k, b in chunks if k == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)  # filter 0 on every row
    return rows[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (300, 250)])
def test_png_round_trip(mcpt, tmp_path, w, h):
    rng = np.random.default_rng(w * h)
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    p = str(tmp_path / "out.png")
    mcpt.write_png(p, img)
    assert np.array_equal(_read_png(p), img)  # 300x250 spans several stored blocks


def test_load_ppm(mcpt, tmp_path):
    """loadPpmAsRgba8 (src/path_tracer.cu:236-294): P6, '#' comment lines in
    the header, maxval 255, alpha 255 appended."""
    px = bytes([10, 20, 30, 40, 50, 60, 70, 80, 90, 200, 210, 220])
    p = tmp_path / "a.ppm"
    p.write_bytes(b"P6\n# made for a test\n2 2\n# another\n255\n" + px)
    img = mcpt.load_ppm(str(p))
    assert img.shape == (2, 2, 4)
    assert np.array_equal(img[..., :3].reshape(-1), np.frombuffer(px, np.uint8)) and np.all(img[..., 3] == 255)
    bad = tmp_path / "b.ppm"
    bad.write_bytes(b"P6\n2 2\n65535\n" + px)
    with pytest.raises(mcpt.McptError):
        mcpt.load_ppm(str(bad))
    short = tmp_path / "c.ppm"
    short.write_bytes(b"P6 2 2 255\n" + px[:7])
    with pytest.raises(mcpt.McptError):
        mcpt.load_ppm(str(short))
