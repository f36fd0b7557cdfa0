"""mcpt_load_png (the PNG textures the reference reads with stb_image,
src/image_io.cpp:30-39) against files written here from the PNG
specification by an independent encoder: every colour type and bit depth,
each of the five row filters, Adam7 interlacing, palettes with and without
tRNS alpha, grey and RGB tRNS keys, image data split over several IDAT
chunks, ancillary chunks skipped; corrupt files (CRC, truncation, an
unknown critical chunk, a bad filter byte, a palette index past the palette,
too little image data) are refused. An OBJ whose map_Kd names a PNG, with
no .ppm beside it, gets the image as its texture. CPU only.
"""
import struct
import zlib

import numpy as np
import pytest

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    """one pass row of samples (w, chans) as bytes, big-endian, MSB first"""
    flat = samples.reshape(-1).astype(np.int64)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in flat)
    if depth == 8:
        return bytes(int(v) for v in flat)
    per = 8 // depth
    out = bytearray((len(flat) + per - 1) // per)
    for i, v in enumerate(flat):
        out[i // per] |= int(v) << (8 - depth - (i % per) * depth)
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter(ftype, raw, prev, bpp):
    out = bytearray(len(raw))
    for i in range(len(raw)):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ftype]
        out[i] = (raw[i] - pred) & 255
    return bytes([ftype]) + bytes(out)


def encode(samples, ctype, depth, interlace=False, plte=None, trns=None, idat_parts=1, extra=()):
    h, w = samples.shape[:2]
    chans = CHANS[ctype]
    bpp = max(1, chans * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    stream = b""
    row_no = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = bytes(len(_pack_row(sub[0], depth)))
        for r in range(sub.shape[0]):
            raw = _pack_row(sub[r], depth)
            stream += _filter(row_no % 5, raw, prev, bpp)
            prev = raw
            row_no += 1
    z = zlib.compress(stream, 9)
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    for c in extra:
        png += c
    if plte is not None:
        png += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        png += _chunk(b"tRNS", trns)
    step = (len(z) + idat_parts - 1) // idat_parts
    for k in range(idat_parts):
        png += _chunk(b"IDAT", z[k * step:(k + 1) * step])
    return png + _chunk(b"IEND", b"")


def expected(samples, ctype, depth, plte=None, trns=None):
    s = samples.astype(np.int64)
    h, w = s.shape[:2]
    to8 = (lambda v: v >> 8) if depth == 16 else (lambda v: v) if depth == 8 else (lambda v: v * (255 // (2 ** depth - 1)))
    out = np.zeros((h, w, 4), np.uint8)
    if ctype == 0:
        out[..., :3] = to8(s[..., :1])
        out[..., 3] = 255
        if trns is not None:
            out[..., 3][s[..., 0] == struct.unpack(">H", trns)[0]] = 0
    elif ctype == 2:
        out[..., :3] = to8(s)
        out[..., 3] = 255
        if trns is not None:
            key = np.array(struct.unpack(">HHH", trns))
            out[..., 3][np.all(s == key, axis=-1)] = 0
    elif ctype == 3:
        pal = np.asarray(plte, np.uint8)
        out[..., :3] = pal[s[..., 0]]
        alpha = np.full(len(pal), 255, np.uint8)
        if trns is not None:
            alpha[:len(trns)] = np.frombuffer(trns, np.uint8)
        out[..., 3] = alpha[s[..., 0]]
    elif ctype == 4:
        out[..., :3] = to8(s[..., :1])
        out[..., 3] = to8(s[..., 1])
    else:
        out[...] = to8(s)
    return out


CASES = [(0, d) for d in (1, 2, 4, 8, 16)] + [(2, 8), (2, 16)] + [(3, d) for d in (1, 2, 4, 8)] + \
        [(4, 8), (4, 16), (6, 8), (6, 16)]


def _samples(rng, ctype, depth, h, w, n_pal=None):
    hi = n_pal if ctype == 3 else 2 ** depth
    return rng.integers(0, hi, (h, w, CHANS[ctype]))


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth", CASES)
def test_png_round_trip(mcpt, tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth * 2 + interlace)
    for h, w in ((1, 1), (7, 13), (17, 9), (3, 33)):
        plte = trns = None
        n_pal = None
        if ctype == 3:
            n_pal = min(2 ** depth, 200)
            plte = rng.integers(0, 256, (n_pal, 3))
            trns = bytes(rng.integers(0, 256, n_pal // 2).astype(np.uint8))
        s = _samples(rng, ctype, depth, h, w, n_pal)
        if ctype in (0, 2) and depth >= 2:  # a tRNS key that some pixels match
            key = s[0, 0]
            trns = b"".join(struct.pack(">H", int(v)) for v in key)
        p = tmp_path / f"t{ctype}_{depth}_{h}x{w}.png"
        p.write_bytes(encode(s, ctype, depth, interlace, plte, trns, idat_parts=3,
                             extra=[_chunk(b"tEXt", b"Comment\x00written by the test")]))
        got = mcpt.load_png(str(p))
        assert got.shape == (h, w, 4)
        assert np.array_equal(got, expected(s, ctype, depth, plte, trns)), (h, w)


def _good(rng):
    s = rng.integers(0, 256, (5, 6, 3))
    return s, encode(s, 2, 8)


def test_png_refuses_corrupt_files(mcpt, tmp_path):
    rng = np.random.default_rng(5)
    s, good = _good(rng)
    p = tmp_path / "good.png"
    p.write_bytes(good)
    assert np.array_equal(mcpt.load_png(str(p)), expected(s, 2, 8))
    bad = {
        "crc": good[:40] + bytes([good[40] ^ 1]) + good[41:],
        "truncated": good[:len(good) - 20],
        "signature": b"\x89PNX" + good[4:],
        "critical": good[:33] + _chunk(b"ABCD", b"xx") + good[33:],
        "no_ihdr": good[:8] + good[33:],
    }
    # a row whose filter byte is 7
    stream = b"".join(b"\x07" + bytes(18) for _ in range(5))
    hdr = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", 6, 5, 8, 2, 0, 0, 0))
    bad["filter"] = hdr + _chunk(b"IDAT", zlib.compress(stream)) + _chunk(b"IEND", b"")
    bad["short_data"] = hdr + _chunk(b"IDAT", zlib.compress(stream[:-1])) + _chunk(b"IEND", b"")
    # a palette of 2 colours and a pixel naming colour 3
    idx = np.zeros((2, 2, 1), np.int64)
    idx[1, 1] = 3
    bad["palette_index"] = encode(idx, 3, 2, plte=[[0, 0, 0], [255, 255, 255]])
    bad["bad_depth"] = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 2, 16, 3, 0, 0, 0)) + \
        _chunk(b"IEND", b"")
    for name, data in bad.items():
        q = tmp_path / f"{name}.png"
        q.write_bytes(data)
        with pytest.raises(mcpt.McptError):
            mcpt.load_png(str(q))
    with pytest.raises(mcpt.McptError):
        mcpt.load_png(str(tmp_path / "missing.png"))


def test_obj_map_kd_png(mcpt, tmp_path):
    rng = np.random.default_rng(6)
    s = rng.integers(0, 256, (6, 10, 4))
    (tmp_path / "tex").mkdir()
    (tmp_path / "tex" / "wall.png").write_bytes(encode(s, 6, 8, interlace=True))
    (tmp_path / "m.mtl").write_text("newmtl w\nKd 1 1 1\nmap_Kd tex\\wall.png\n")
    (tmp_path / "t.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl w\nf 1 2 3\n")
    hot, cold, mats, st, images = mcpt.load_obj(str(tmp_path / "t.obj"), load_textures=True)
    assert st["textures_loaded"] == 1 and st["missing_textures"] == 0
    assert np.array_equal(images[0], expected(s, 6, 8))
