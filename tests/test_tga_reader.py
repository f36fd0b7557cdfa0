"""mcpt_load_tga against TGA files written here by an independent encoder:
true-colour 16/24/32-bit, grey and colour-mapped (24- and 32-bit palette
entries with a nonzero first index) images, raw and run-length encoded,
bottom-first and top-first rows; run packets that cross scanlines; corrupt
files (short pixel data, a run past the last pixel, a colour index outside
the map, an unsupported type) refused; an OBJ whose map_Kd names a TGA.
CPU only.
"""
import struct

import numpy as np
import pytest


def _px_bytes(kind, v, bits):
    if kind == 3 or kind == 1:
        return bytes([int(v)])
    if bits == 16:
        return struct.pack("<H", int(v))
    r, g, b, a = (int(x) for x in v)
    return bytes([b, g, r] + ([a] if bits == 32 else []))


def encode(values, kind, bits, rle=False, top_first=False, cmap=None, cmfirst=0, cmbits=24):
    """values: (h, w) ints for grey / indices / 16-bit words, (h, w, 4) RGBA otherwise"""
    h, w = values.shape[:2]
    rows = values if top_first else values[::-1]
    flat = [rows[r, c] for r in range(h) for c in range(w)]
    itype = kind + (8 if rle else 0)
    cm = b""
    if cmap is not None:
        for e in cmap:
            cm += bytes([e[2], e[1], e[0]] + ([e[3]] if cmbits == 32 else []))
    hdr = struct.pack("<BBBHHBHHHHBB", 3, 1 if cmap is not None else 0, itype, cmfirst,
                      0 if cmap is None else len(cmap), cmbits if cmap is not None else 0, 0, 0, w, h, bits,
                      0x20 if top_first else 0)
    body = b""
    if not rle:
        body = b"".join(_px_bytes(kind, v, bits) for v in flat)
    else:  # alternate run and literal packets, crossing scanlines
        i, toggle = 0, True
        while i < len(flat):
            n = min(len(flat) - i, 5 if toggle else 3)
            if toggle:  # a run of the first pixel's value
                body += bytes([0x80 | (n - 1)]) + _px_bytes(kind, flat[i], bits)
                for k in range(n):
                    flat[i + k] = flat[i]
            else:
                body += bytes([n - 1]) + b"".join(_px_bytes(kind, flat[i + k], bits) for k in range(n))
            i += n
            toggle = not toggle
        rows = np.array(flat, dtype=object).reshape((h, w) + values.shape[2:])
        values = rows if top_first else rows[::-1]
    return hdr + b"ABC" + cm + body, values


def expected(values, kind, bits, cmap=None, cmfirst=0, cmbits=24):
    h, w = values.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    for r in range(h):
        for c in range(w):
            v = values[r, c]
            if kind == 3:
                out[r, c] = (v, v, v, 255)
            elif kind == 1:
                e = cmap[int(v) - cmfirst]
                out[r, c] = (e[0], e[1], e[2], e[3] if cmbits == 32 else 255)
            elif bits == 16:
                v = int(v)
                out[r, c] = (((v >> 10) & 31) * 255 // 31, ((v >> 5) & 31) * 255 // 31, (v & 31) * 255 // 31, 255)
            else:
                out[r, c] = (v[0], v[1], v[2], v[3] if bits == 32 else 255)
    return out


CASES = [(2, 24), (2, 32), (2, 16), (3, 8), (1, 8)]


@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("top_first", [False, True])
@pytest.mark.parametrize("kind,bits", CASES)
def test_tga_round_trip(mcpt, tmp_path, kind, bits, rle, top_first):
    rng = np.random.default_rng(kind * 100 + bits + 7 * rle + 3 * top_first)
    for h, w in ((1, 1), (5, 7), (9, 4)):
        cmap, cmfirst, cmbits = None, 0, 24
        if kind == 1:
            cmfirst, cmbits = 3, 32 if top_first else 24
            cmap = rng.integers(0, 256, (20, 4))
            vals = rng.integers(cmfirst, cmfirst + 20, (h, w))
        elif kind == 3:
            vals = rng.integers(0, 256, (h, w))
        elif bits == 16:
            vals = rng.integers(0, 65536, (h, w))
        else:
            vals = rng.integers(0, 256, (h, w, 4))
        data, vals = encode(vals, kind, bits, rle, top_first, cmap, cmfirst, cmbits)
        p = tmp_path / f"t{kind}_{bits}_{h}x{w}.tga"
        p.write_bytes(data)
        got = mcpt.load_tga(str(p))
        assert np.array_equal(got, expected(vals, kind, bits, cmap, cmfirst, cmbits)), (h, w)


def test_tga_refuses_corrupt_files(mcpt, tmp_path):
    rng = np.random.default_rng(1)
    vals = rng.integers(0, 256, (4, 4, 4))
    good, _ = encode(vals, 2, 24)
    hdr = good[:21]
    bad = {"short": good[:-1], "header": good[:10], "type": good[:2] + b"\x05" + good[3:]}
    # a run packet of 32 pixels in a 16-pixel image
    bad["run_past_end"] = hdr + bytes([0x80 | 31]) + b"\x00\x00\x00"
    # a colour index below the map's first entry
    cmap = rng.integers(0, 256, (4, 4))
    cm, _ = encode(np.full((2, 2), 5), 1, 8, cmap=cmap, cmfirst=5)
    idx_at = len(cm) - 4
    bad["index"] = cm[:idx_at] + b"\x04" + cm[idx_at + 1:]
    for name, data in bad.items():
        q = tmp_path / f"{name}.tga"
        q.write_bytes(data)
        with pytest.raises(mcpt.McptError):
            mcpt.load_tga(str(q))
    ok = tmp_path / "good.tga"
    ok.write_bytes(good)
    assert mcpt.load_tga(str(ok)).shape == (4, 4, 4)


def test_obj_map_kd_tga(mcpt, tmp_path):
    rng = np.random.default_rng(2)
    vals = rng.integers(0, 256, (6, 5, 4))
    data, vals = encode(vals, 2, 32, rle=True)
    (tmp_path / "brick.TGA").write_bytes(data)
    (tmp_path / "m.mtl").write_text("newmtl b\nKd 1 1 1\nmap_Kd brick.TGA\n")
    (tmp_path / "t.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl b\nf 1 2 3\n")
    _, _, mats, st, images = mcpt.load_obj(str(tmp_path / "t.obj"), load_textures=True)
    assert st["textures_loaded"] == 1 and mats["texture"][-1] == 1
    assert np.array_equal(images[0], expected(vals, 2, 32))
