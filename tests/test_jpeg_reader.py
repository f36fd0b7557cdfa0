"""mcpt_load_jpeg (baseline JPEG textures) against an independent decoder,
libjpeg-turbo through Pillow, on files Pillow writes: RGB with 4:2:0, 4:2:2
and 4:4:4 chroma, grey, odd sizes, several qualities, restart intervals;
against the reference's own texture pair (models/earth.jpeg and the
earth.ppm decoded from it); progressive and damaged files refused; an OBJ
whose map_Kd names a JPEG. CPU only.

Bar: two decoders of the same file differ by IDCT and upsampling rounding
only, so the images agree within a few 8-bit levels (mean difference below
half a level); a wrong Huffman table, zig-zag order, quantization, sampling
layout or colour transform would move most pixels by tens of levels.
"""
import io
import os

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")


def _photo(h, w, seed):
    """smooth colour gradients plus texture, so every coefficient band is used"""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.stack([127 + 100 * np.sin(x / 7 + seed), 127 + 100 * np.cos(y / 5), 127 + 60 * np.sin((x + y) / 9)], -1)
    img += rng.normal(0, 12, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


def _write(path, arr, **kw):
    PIL.fromarray(arr).save(path, "JPEG", **kw)
    return str(path)


def _check(mcpt, path, grey=False):
    got = mcpt.load_jpeg(path)
    want = np.asarray(PIL.open(path).convert("RGBA"))
    assert got.shape == want.shape
    assert np.all(got[..., 3] == 255)
    d = np.abs(got[..., :3].astype(int) - want[..., :3].astype(int))
    assert d.mean() < 0.5 and d.max() <= 4, (d.mean(), d.max())
    if grey:
        assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    return got


@pytest.mark.parametrize("subsampling", [0, 1, 2])  # 4:4:4, 4:2:2, 4:2:0
@pytest.mark.parametrize("h,w", [(64, 64), (37, 23), (67, 129)])
@pytest.mark.parametrize("quality", [95, 50])
def test_rgb(mcpt, tmp_path, subsampling, h, w, quality):
    arr = _photo(h, w, subsampling + h)
    _check(mcpt, _write(tmp_path / "a.jpg", arr, quality=quality, subsampling=subsampling))


@pytest.mark.parametrize("h,w", [(16, 16), (45, 30)])
def test_grey(mcpt, tmp_path, h, w):
    arr = _photo(h, w, 3)[..., 0]
    _check(mcpt, _write(tmp_path / "g.jpg", arr, quality=90), grey=True)


@pytest.mark.parametrize("blocks", [1, 3, 7])
def test_restart_intervals(mcpt, tmp_path, blocks):
    arr = _photo(50, 70, blocks)
    p = _write(tmp_path / "r.jpg", arr, quality=85, subsampling=2, restart_marker_blocks=blocks)
    data = open(p, "rb").read()
    assert b"\xff\xdd" in data and b"\xff\xd0" in data
    _check(mcpt, p)


def test_size_query(mcpt, tmp_path):
    import ctypes as C
    p = _write(tmp_path / "s.jpg", _photo(19, 33, 1))
    w, h = C.c_int32(), C.c_int32()
    assert mcpt._lib.mcpt_load_jpeg(os.fsencode(p), C.byref(w), C.byref(h), None) == 0
    assert (w.value, h.value) == (33, 19)


def test_refused(mcpt, tmp_path):
    arr = _photo(32, 32, 5)
    p = _write(tmp_path / "p.jpg", arr, progressive=True)
    with pytest.raises(mcpt.McptError) as e:
        mcpt.load_jpeg(p)
    assert "progressive" in str(e.value)
    good = open(_write(tmp_path / "b.jpg", arr), "rb").read()
    for cut in (2, 20, len(good) // 3):
        (tmp_path / "t.jpg").write_bytes(good[:cut])
        with pytest.raises(mcpt.McptError):
            mcpt.load_jpeg(str(tmp_path / "t.jpg"))
    # a Huffman table whose code lengths overflow their bit lengths
    dht = good.index(b"\xff\xc4")
    bad = bytearray(good)
    bad[dht + 5] = 5  # five codes of length 1
    (tmp_path / "h.jpg").write_bytes(bytes(bad))
    with pytest.raises(mcpt.McptError):
        mcpt.load_jpeg(str(tmp_path / "h.jpg"))
    (tmp_path / "n.jpg").write_bytes(b"not a jpeg at all")
    with pytest.raises(mcpt.McptError):
        mcpt.load_jpeg(str(tmp_path / "n.jpg"))


EARTH = "/root/reference/models/earth.jpeg"


@pytest.mark.skipif(not os.path.exists(EARTH), reason="the reference's texture asset is not on this machine")
def test_reference_earth_jpeg(mcpt):
    """The reference's JPEG texture (baseline, 632x316, 4:2:0) against the
    PPM the reference ships beside it, decoded from the same file."""
    got = mcpt.load_jpeg(EARTH)
    want = mcpt.load_ppm(EARTH[:-5] + ".ppm")
    assert got.shape == want.shape == (316, 632, 4)
    d = np.abs(got[..., :3].astype(int) - want[..., :3].astype(int))
    assert d.mean() < 0.1 and d.max() <= 3 and (d == 0).mean() > 0.95, (d.mean(), d.max(), (d == 0).mean())


def test_obj_map_kd_jpeg(mcpt, tmp_path):
    arr = _photo(12, 10, 9)
    p = _write(tmp_path / "wood.jpg", arr, quality=90)
    (tmp_path / "m.mtl").write_text("newmtl b\nKd 1 1 1\nmap_Kd wood.jpg\n")
    (tmp_path / "t.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl b\nf 1 2 3\n")
    _, _, mats, st, images = mcpt.load_obj(str(tmp_path / "t.obj"), load_textures=True)
    assert st["textures_loaded"] == 1 and mats["texture"][-1] == 1
    assert np.array_equal(images[0], mcpt.load_jpeg(p))
