"""Image files a scene reads (host code, no GPU call): PFM and Radiance HDR
environment maps and PPM textures.

PFM follows loadPfmAsRgba32f (src/path_tracer.cu:335-368); the Radiance
decode follows the format the reference reads with stb_image and converts in
scripts/hdr2pfm.py:14-61. The checks are known answers written here: a PFM
written bottom row first loads back top row first bit for bit, and every
RGBE texel decodes to mantissa * 2^(e - 136) exactly (0 for e = 0), for
run-length and flat rows alike. The reference's own texture asset
(models/earth.ppm) is compared with an independent parse when present."""
import os
import struct
import zlib

import numpy as np
import pytest


def _pfm(path, rgba, scale=b"-1.0", header=None):
    h, w = rgba.shape[:2]
    head = header if header is not None else b"PF\n%d %d\n%s\n" % (w, h, scale)
    path.write_bytes(head + np.ascontiguousarray(rgba[::-1, :, :3], "<f4").tobytes())
    return str(path)


def test_pfm_round_trip(mcpt, tmp_path):
    sky = mcpt.gen_sky(64, 32)
    got = mcpt.load_pfm(_pfm(tmp_path / "sky.pfm", sky))
    assert got.shape == (32, 64, 4) and np.array_equal(got[..., :3], sky[..., :3]) and np.all(got[..., 3] == 1)
    assert np.array_equal(mcpt.load_env(str(tmp_path / "sky.pfm")), got)
    # a one-line header, one separator byte before the pixels
    small = np.arange(3 * 2 * 4, dtype=np.float32).reshape(2, 3, 4)
    got = mcpt.load_pfm(_pfm(tmp_path / "s.pfm", small, header=b"PF 3 2 -1 "))
    assert np.array_equal(got[..., :3], small[..., :3])


@pytest.mark.parametrize("case", ["big_endian", "grey", "short", "zero_width"])
def test_pfm_refused(mcpt, tmp_path, case):
    img = np.ones((4, 5, 4), np.float32)
    if case == "big_endian":
        p = _pfm(tmp_path / "a.pfm", img, scale=b"1.0")
    elif case == "grey":
        p = _pfm(tmp_path / "a.pfm", img, header=b"Pf\n5 4\n-1.0\n")
    elif case == "short":
        p = _pfm(tmp_path / "a.pfm", img)
        data = open(p, "rb").read()
        open(p, "wb").write(data[:-4])
    else:
        p = _pfm(tmp_path / "a.pfm", img, header=b"PF\n0 4\n-1.0\n")
    with pytest.raises(mcpt.McptError):
        mcpt.load_pfm(p)


def _rle_plane(v):
    """New-style Radiance run-length encoding of one byte plane."""
    out, i, n = bytearray(), 0, len(v)
    while i < n:
        j = i
        while j < n and j - i < 127 and v[j] == v[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), v[i]])
            i = j
            continue
        j = i + 1
        while j < n and j - i < 128 and not (j + 2 < n and v[j] == v[j + 1] == v[j + 2]):
            j += 1
        out += bytes([j - i]) + bytes(v[i:j])
        i = j
    return bytes(out)


def _hdr(path, rgbe, rle_rows=(), header=b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n", res=None):
    h, w = rgbe.shape[:2]
    body = bytearray(header + (res if res is not None else b"-Y %d +X %d\n" % (h, w)))
    for y in range(h):
        if y in rle_rows:
            body += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                body += _rle_plane(rgbe[y, :, c].tolist())
        else:
            body += rgbe[y].tobytes()
    path.write_bytes(bytes(body))
    return str(path)


def _decode(rgbe):
    e = rgbe[..., 3].astype(np.int64)
    s = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return rgbe[..., :3].astype(np.float64) * s[..., None]


@pytest.mark.parametrize("w", [5, 40, 300])
def test_hdr_decode(mcpt, tmp_path, w):
    rng = np.random.default_rng(w)
    h = 6
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[..., 3] = rng.choice([0, 120, 128, 129, 140, 255], (h, w))
    rgbe[1, 3:20] = rgbe[1, 2]  # runs for the run-length rows
    rgbe[2, :, 3] = 130
    rows = (1, 2, 4) if w >= 8 else ()
    got = mcpt.load_hdr(_hdr(tmp_path / "a.hdr", rgbe, rle_rows=rows))
    assert got.shape == (h, w, 4) and np.all(got[..., 3] == 1)
    want = _decode(rgbe)
    assert np.array_equal(got[..., :3].astype(np.float64), want)  # exact: small integer times a power of two
    assert np.array_equal(mcpt.load_env(str(tmp_path / "a.hdr")), got)


@pytest.mark.parametrize("case", ["signature", "orientation", "xyze", "zero_count", "overrun", "short"])
def test_hdr_refused(mcpt, tmp_path, case):
    rgbe = np.full((3, 16, 4), 7, np.uint8)
    p = tmp_path / "a.hdr"
    if case == "signature":
        _hdr(p, rgbe, header=b"RADIANCE\n\n")
    elif case == "orientation":
        _hdr(p, rgbe, res=b"+Y 3 +X 16\n")
    elif case == "xyze":
        _hdr(p, rgbe, header=b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n")
    else:
        _hdr(p, rgbe, rle_rows=(0,))
        data = bytearray(p.read_bytes())
        at = data.index(b"+X 16\n") + 6 + 4  # first count byte of row 0's first plane
        if case == "zero_count":
            data[at] = 0
        elif case == "overrun":
            data[at] = 128 + 17  # a run past the row's end
        else:
            data = data[:-5]
        p.write_bytes(bytes(data))
    with pytest.raises(mcpt.McptError):
        mcpt.load_hdr(str(p))


def test_env_extension(mcpt, tmp_path):
    with pytest.raises(ValueError):
        mcpt.load_env(str(tmp_path / "sky.tga"))


# ------------------------------------------------------------- OpenEXR
# A writer of the published OpenEXR scanline layout, for the reader's known
# answers: magic, version 2, the header attributes, the chunk offset table,
# and per chunk its first row, its size and the rows (channels in name
# order); ZIP / ZIPS blocks are the rows with their bytes split into even
# and odd halves, delta-coded (+128) and deflated, RLE blocks the same
# delta-coded bytes in signed-count runs.
EXR_TYPES = {0: np.uint32, 1: np.float16, 2: np.float32}


def _exr_predict(raw):
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 255
    return d.astype(np.uint8).tobytes()


def _exr_rle(b):
    out, i = bytearray(), 0
    while i < len(b):
        j = i
        while j < len(b) and j - i < 128 and b[j] == b[i]:
            j += 1
        if j - i >= 3:
            out += struct.pack("<bB", j - i - 1, b[i])
            i = j
            continue
        j = i + 1
        while j < len(b) and j - i < 127 and not (j + 2 < len(b) and b[j] == b[j + 1] == b[j + 2]):
            j += 1
        out += struct.pack("<b", -(j - i)) + b[i:j]
        i = j
    return bytes(out)


def _exr(path, planes, comp, ymin=0, decreasing=False):
    """planes: {name: (type, (H, W) array)}"""
    names = sorted(planes)
    H, W = planes[names[0]][1].shape

    def attr(n, t, v):
        return n.encode() + b"\0" + t.encode() + b"\0" + struct.pack("<i", len(v)) + v

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", planes[n][0], 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, ymin, W - 1, ymin + H - 1)
    hdr = (attr("channels", "chlist", chl) + attr("compression", "compression", bytes([comp])) +
           attr("dataWindow", "box2i", box) + attr("displayWindow", "box2i", box) +
           attr("lineOrder", "lineOrder", bytes([1 if decreasing else 0])) +
           attr("pixelAspectRatio", "float", struct.pack("<f", 1)) +
           attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
           attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    lines = 16 if comp == 3 else 1
    chunks = []
    for y0 in range(0, H, lines):
        rows = range(y0, min(H, y0 + lines))
        raw = b"".join(np.ascontiguousarray(planes[n][1][y], EXR_TYPES[planes[n][0]]).tobytes()
                       for y in rows for n in names)
        if comp in (2, 3):
            pay = zlib.compress(_exr_predict(raw))
        elif comp == 1:
            pay = _exr_rle(_exr_predict(raw))
        else:
            pay = raw
        if len(pay) >= len(raw):
            pay = raw
        chunks.append((y0, struct.pack("<ii", ymin + y0, len(pay)) + pay))
    order = chunks[::-1] if decreasing else chunks
    base = 8 + len(hdr) + 8 * len(chunks)
    offs, body, at = {}, b"", base
    for y0, c in order:
        offs[y0] = at
        body += c
        at += len(c)
    table = b"".join(struct.pack("<Q", offs[y0]) for y0, _ in chunks)
    path.write_bytes(struct.pack("<ii", 20000630, 2) + hdr + table + body)
    return str(path)


@pytest.mark.parametrize("comp", [0, 1, 2, 3])
def test_exr_decode(mcpt, tmp_path, comp):
    rng = np.random.default_rng(comp)
    H, W = 37, 23
    r = rng.uniform(0, 50, (H, W)).astype(np.float16)
    r[0, :4] = [np.float16(6e-8), np.float16(np.inf), np.float16(65504), np.float16(-2.5)]  # subnormal, inf, max
    g = rng.uniform(0, 5, (H, W)).astype(np.float32)
    b = np.repeat(rng.uniform(0, 1, (H, 1)), W, 1).astype(np.float16)  # runs for RLE
    a = rng.integers(0, 1000, (H, W)).astype(np.uint32)
    planes = {"R": (1, r), "G": (2, g), "B": (1, b), "A": (0, a)}
    got = mcpt.load_exr(_exr(tmp_path / "a.exr", planes, comp, ymin=-5, decreasing=comp == 2))
    assert got.shape == (H, W, 4)
    assert np.array_equal(got[..., 0], r.astype(np.float32))
    assert np.array_equal(got[..., 1], g) and np.array_equal(got[..., 2], b.astype(np.float32))
    assert np.array_equal(got[..., 3], a.astype(np.float32))
    assert np.array_equal(mcpt.load_env(str(tmp_path / "a.exr")), got)


def test_exr_channels_and_refusals(mcpt, tmp_path):
    y = np.arange(12, dtype=np.float16).reshape(3, 4)
    got = mcpt.load_exr(_exr(tmp_path / "y.exr", {"Y": (1, y)}, 3))
    assert np.array_equal(got[..., 0], y.astype(np.float32)) and np.array_equal(got[..., 2], got[..., 0])
    assert np.all(got[..., 3] == 1)
    rgb = {"R": (2, np.ones((2, 2), np.float32)), "G": (2, np.zeros((2, 2), np.float32)),
           "B": (2, np.ones((2, 2), np.float32))}
    assert np.all(mcpt.load_exr(_exr(tmp_path / "rgb.exr", rgb, 0))[..., 3] == 1)  # no A: 1
    p = _exr(tmp_path / "piz.exr", rgb, 0)
    data = bytearray(open(p, "rb").read())
    at = data.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    data[at] = 4  # PIZ
    open(p, "wb").write(bytes(data))
    with pytest.raises(mcpt.McptError):
        mcpt.load_exr(p)
    with pytest.raises(mcpt.McptError):
        mcpt.load_exr(_pfm(tmp_path / "not.exr", np.ones((2, 2, 4), np.float32)))


@pytest.mark.parametrize("off,size", [(2**64 - 8, None), (2**64 - 1, None), (2**63, None), (None, 2**32 - 1),
                                      (None, 2**32 - 16)])
def test_exr_crafted_chunk_table(mcpt, tmp_path, off, size):
    """Chunk offsets near 2^64 and chunk sizes near 2^32 are refused as such,
    without wrapping past the bounds test."""
    rgb = {"R": (2, np.ones((2, 2), np.float32)), "G": (2, np.zeros((2, 2), np.float32)),
           "B": (2, np.ones((2, 2), np.float32))}
    p = _exr(tmp_path / "crafted.exr", rgb, 0)
    data = bytearray(open(p, "rb").read())
    table = data.index(b"\0", data.index(b"screenWindowWidth\0float\0") + 30) + 1  # after the header's last byte
    first = struct.unpack_from("<Q", data, table)[0]
    if off is not None:
        struct.pack_into("<Q", data, table, off)
    else:
        struct.pack_into("<I", data, first + 4, size)
    open(p, "wb").write(bytes(data))
    with pytest.raises(mcpt.McptError) as e:
        mcpt.load_exr(p)
    assert ("chunk offset" if off is not None else "truncated chunk") in str(e.value)


EARTH = "/root/reference/models/earth.ppm"


@pytest.mark.skipif(not os.path.exists(EARTH), reason="the reference's texture asset is not on this machine")
def test_reference_earth_ppm(mcpt):
    """The reference's own texture (loaded by its earth scene,
    src/path_tracer.cu:1782-1793) against a parse written here."""
    data = open(EARTH, "rb").read()
    tok, at = [], 2
    assert data[:2] == b"P6"
    while len(tok) < 3:
        while data[at:at + 1].isspace():
            at += 1
        if data[at:at + 1] == b"#":
            at = data.index(b"\n", at)
            continue
        end = at
        while data[end:end + 1].isdigit():
            end += 1
        tok.append(int(data[at:end]))
        at = end
    at += 1
    w, h, maxv = tok
    assert maxv == 255
    want = np.frombuffer(data, np.uint8, w * h * 3, at).reshape(h, w, 3)
    img = mcpt.load_ppm(EARTH)
    assert img.shape == (h, w, 4) and np.array_equal(img[..., :3], want) and np.all(img[..., 3] == 255)
