"""The NanoVDB reader (mcpt_load_nvdb, csrc/nvdb.cpp) on the CPU.

The reference reads volumes with NanoVDB's io::readGrid (src/volume_io.cpp:
45-53) and samples the sparse float grid trilinearly in index space
(src/path_tracer.cu:784-789). NanoVDB (32.7, OpenVDB 12) is absent from
this image and the reference holds no .nvdb file, so the file format is
parity unpinned: the files here come from `_write_nvdb` below, which lays
out NanoVDB's records from the published format as the reader does, and
only pin the reader's walk of that layout. What they do pin:

  * the densified grid against the voxel values the writer was given,
    through every kind of node: 8^3 leaves (value order x-major, then y,
    then z), tile values of 16^3 and 32^3 internal nodes, a root tile
    without a child at a negative block key, children that the root's index
    box clips, and child masks that send a slot to a child or to its value;
  * world coordinates: the dense grid's trilinear sample (the oracle's
    grid_density) at random world points equals the sparse grid sampled the
    NanoVDB way (world to index through the inverse of the map, then
    trilinear over the voxel values, background outside);
  * raw and ZIP-compressed grids, the first float grid of a file whose first
    grid is of another type;
  * refusals: a bad magic, truncation, BLOSC, rotation in the map, a
    non-zero background.
"""
import math
import struct
import zlib

import numpy as np
import pytest

MAGIC_FILE = 0x324244566f6e614e
MAGIC_GRID = 0x314244566f6e614e
VERSION = (32 << 21) | (7 << 10)


def _key(i, j, k):
    return ((k & 0xFFFFFFFF) >> 12) | (((j & 0xFFFFFFFF) >> 12) << 21) | (((i & 0xFFFFFFFF) >> 12) << 42)


def _pad32(b):
    return b + b"\0" * (-len(b) % 32)


def _mask(bits, n):
    words = [0] * (n // 64)
    for b in bits:
        words[b >> 6] |= 1 << (b & 63)
    return struct.pack(f"<{len(words)}Q", *words)


def _leaf(origin, values):
    """values: (8, 8, 8) float indexed [x][y][z]"""
    head = struct.pack("<3i3BB", *origin, 7, 7, 7, 0) + _mask(range(512), 512)
    head += struct.pack("<4f", float(values.min()), float(values.max()), 0.0, 0.0)
    assert len(head) == 96
    return head + np.asarray(values, np.float32).reshape(512).tobytes()


def _internal(log2, origin, children, tiles):
    """children: {slot: bytes of the child}; tiles: {slot: value}. Returns the
    node's bytes followed by its children's (offsets from the node)."""
    n = 1 << (3 * log2)
    head = struct.pack("<6iQ", *origin, *origin, 0)
    head += _mask(list(tiles), n) + _mask(list(children), n) + struct.pack("<4f", 0, 1, 0, 0)
    table_at = len(_pad32(head))
    body = b""
    entries = [struct.pack("<f4x", 0.0)] * n
    for s, v in tiles.items():
        entries[s] = struct.pack("<f4x", v)
    child_at = table_at + 8 * n
    for s, c in sorted(children.items()):
        entries[s] = struct.pack("<q", child_at + len(body))
        body += _pad32(c)
    return _pad32(head) + b"".join(entries) + body


def _slot(log2, child_total, i, j, k):
    m = (1 << (log2 + child_total)) - 1
    return (((i & m) >> child_total) << (2 * log2)) | (((j & m) >> child_total) << log2) | ((k & m) >> child_total)


def _grid_bytes(root_tiles, bbox, mat, vec, background=0.0, maximum=1.0, grid_type=1):
    """root_tiles: list of (origin, child bytes or None, value)"""
    root = struct.pack("<6iIfffff", *bbox[0], *bbox[1], len(root_tiles), background, 0.0, maximum, 0.0, 0.0)
    root = _pad32(root)
    assert len(root) == 64
    tiles, kids = b"", b""
    child_base = 64 + 32 * len(root_tiles)
    for origin, child, value in root_tiles:
        off = 0
        if child is not None:
            off = child_base + len(kids)
            kids += _pad32(child)
        tiles += _pad32(struct.pack("<QqIf", _key(*origin), off, 1, value))
    root_bytes = root + tiles + kids
    tree = struct.pack("<4q3I3IQ", 0, 0, 0, 64, 0, 0, 0, 0, 0, len(root_tiles), 0)
    assert len(tree) == 64
    inv = np.linalg.inv(np.asarray(mat, np.float64).reshape(3, 3)).reshape(9)
    mp = struct.pack("<9f9f3ff", *mat, *inv, *vec, 1.0) + struct.pack("<9d9d3dd", *mat, *inv, *vec, 1.0)
    assert len(mp) == 264
    size = 672 + 64 + len(root_bytes)
    head = struct.pack("<QQIIIIQ", MAGIC_GRID, 0, VERSION, 0, 0, 1, size) + b"density".ljust(256, b"\0") + mp
    head += struct.pack("<6d3dIIqIIQQ", 0, 0, 0, 1, 1, 1, mat[0], mat[4], mat[8], 2, grid_type, 0, 0, 0, 0, 0)
    assert len(head) == 672
    return head + tree + root_bytes


def _write_nvdb(path, grids, codec=0):
    """grids: list of (grid bytes, grid type)"""
    metas, bufs = b"", b""
    for g, gtype in grids:
        stored = g if codec == 0 else struct.pack("<Q", len(zlib.compress(g))) + zlib.compress(g)
        name = b"density\0"
        metas += struct.pack("<QQQQII6d6i3dI4I3IHHI", len(g), len(stored), 0, 0, gtype, 2, *([0.0] * 6),
                             *([0] * 6), 1, 1, 1, len(name), 0, 0, 0, 0, 0, 0, 0, codec, 0, VERSION) + name
        bufs += stored
    with open(path, "wb") as f:
        f.write(struct.pack("<QIHH", MAGIC_FILE, VERSION, len(grids), codec) + metas + bufs)


def _field(i, j, k):
    return 0.125 + 0.01 * i + 0.02 * j + 0.03 * k


def _scene_grid():
    """The test grid: returns (root tiles, value(i, j, k), index bbox)."""
    bbox = ((-8, -3, 0), (37, 21, 23))
    # block (0, 0, 0): an upper node whose first lower node holds two leaves
    # and a tile value; the upper node has one tile value of its own
    leaf_a = np.fromfunction(lambda x, y, z: _field(x, y, z), (8, 8, 8))
    leaf_b = np.fromfunction(lambda x, y, z: _field(x + 8, y, z + 16), (8, 8, 8))
    lower = _internal(4, (0, 0, 0), {_slot(4, 3, 0, 0, 0): _leaf((0, 0, 0), leaf_a),
                                     _slot(4, 3, 8, 0, 16): _leaf((8, 0, 16), leaf_b)},
                      {_slot(4, 3, 16, 8, 0): 0.5, _slot(4, 3, 24, 16, 8): 0.75})
    upper = _internal(5, (0, 0, 0), {_slot(5, 7, 0, 0, 0): lower}, {_slot(5, 7, 0, 0, 128): 2.0})
    root_tiles = [((0, 0, 0), upper, 0.0), ((-4096, 0, 0), None, 0.25)]

    def value(i, j, k):
        if i < 0:
            return 0.25 if 0 <= j < 4096 and 0 <= k < 4096 else 0.0
        if 0 <= i < 8 and 0 <= j < 8 and 0 <= k < 8:
            return _field(i, j, k)
        if 8 <= i < 16 and 0 <= j < 8 and 16 <= k < 24:
            return _field(i, j, k)
        if 16 <= i < 24 and 8 <= j < 16 and 0 <= k < 8:
            return 0.5
        if 24 <= i < 32 and 16 <= j < 24 and 8 <= k < 16:
            return 0.75
        return 0.0
    return root_tiles, value, bbox


MAT = [0.25, 0, 0, 0, 0.5, 0, 0, 0, 0.75]
VEC = [1.0, -2.0, 0.5]


def _expected(value, bbox):
    (x0, y0, z0), (x1, y1, z1) = bbox
    out = np.zeros((z1 - z0 + 1, y1 - y0 + 1, x1 - x0 + 1), np.float32)
    for k in range(z0, z1 + 1):
        for j in range(y0, y1 + 1):
            for i in range(x0, x1 + 1):
                out[k - z0, j - y0, i - x0] = value(i, j, k)
    return out


@pytest.mark.parametrize("codec", [0, 1])
def test_densified_grid_matches_the_voxels(mcpt, tmp_path, codec):
    root_tiles, value, bbox = _scene_grid()
    g = _grid_bytes(root_tiles, bbox, MAT, VEC, maximum=2.0)
    other = _grid_bytes(root_tiles, bbox, MAT, VEC, grid_type=2)  # a double grid comes first
    path = tmp_path / "grid.nvdb"
    _write_nvdb(path, [(other, 2), (g, 1)], codec)
    dens, origin, voxel, mx = mcpt.load_nvdb(path)
    want = _expected(value, bbox)
    assert dens.shape == want.shape
    assert np.array_equal(dens, want)
    assert mx == 2.0
    assert voxel == (0.25, 0.5, 0.75)
    assert origin == tuple(float(np.float32(VEC[c] + MAT[4 * c] * bbox[0][c])) for c in range(3))


def _nanovdb_sample(value, p):
    """NanoVDB's trilinear sampler at world point p: index = inv(mat)(p - vec)"""
    q = [(p[c] - VEC[c]) / MAT[4 * c] for c in range(3)]
    f = [math.floor(x) for x in q]
    u = [q[c] - f[c] for c in range(3)]
    acc = 0.0
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = (u[0] if di else 1 - u[0]) * (u[1] if dj else 1 - u[1]) * (u[2] if dk else 1 - u[2])
                acc += w * value(f[0] + di, f[1] + dj, f[2] + dk)
    return acc


def test_world_sampling_matches_the_sparse_grid(mcpt, oracle, tmp_path):
    root_tiles, value, bbox = _scene_grid()
    path = tmp_path / "grid.nvdb"
    _write_nvdb(path, [(_grid_bytes(root_tiles, bbox, MAT, VEC, maximum=2.0), 1)])
    dens, origin, voxel, _ = mcpt.load_nvdb(path)
    sp = np.zeros(1, oracle.SPHERE_DTYPE)
    sp["radius"] = 1.0
    s = oracle.Scene(np.zeros(0, oracle.HOT_DTYPE), np.zeros(0, oracle.COLD_DTYPE), np.zeros(1, oracle.MAT_DTYPE),
                     spheres=sp)
    s.set_density_grid(dens, origin, voxel)
    rng = np.random.default_rng(3)
    lo = np.array([VEC[c] + MAT[4 * c] * (bbox[0][c] + 0.01) for c in range(3)])
    hi = np.array([VEC[c] + MAT[4 * c] * (bbox[1][c] - 0.01) for c in range(3)])
    pts = rng.uniform(lo, hi, (3000, 3)).astype(np.float32)
    got = s.grid_density(pts)
    want = np.array([_nanovdb_sample(value, p.astype(np.float64)) for p in pts])
    # the dense grid forms (p - origin) / voxel in FP32, NanoVDB inv(mat)(p - vec)
    assert np.max(np.abs(got - want)) < 2e-4
    assert (want > 0).mean() > 0.2


def test_refusals(mcpt, tmp_path):
    root_tiles, _, bbox = _scene_grid()
    good = _grid_bytes(root_tiles, bbox, MAT, VEC)
    cases = {}
    p = tmp_path / "magic.nvdb"
    _write_nvdb(p, [(good, 1)])
    b = bytearray(p.read_bytes())
    b[0] ^= 1
    p.write_bytes(bytes(b))
    cases["magic"] = p
    p = tmp_path / "short.nvdb"
    _write_nvdb(p, [(good, 1)])
    p.write_bytes(p.read_bytes()[:-100])
    cases["truncated"] = p
    p = tmp_path / "blosc.nvdb"
    _write_nvdb(p, [(good, 1)])
    b = bytearray(p.read_bytes())
    struct.pack_into("<H", b, 16 + 168, 2)
    p.write_bytes(bytes(b))
    cases["blosc"] = p
    p = tmp_path / "rot.nvdb"
    _write_nvdb(p, [(_grid_bytes(root_tiles, bbox, [0.5, 0.1, 0, 0, 0.5, 0, 0, 0, 0.5], VEC), 1)])
    cases["rotation"] = p
    p = tmp_path / "bg.nvdb"
    _write_nvdb(p, [(_grid_bytes(root_tiles, bbox, MAT, VEC, background=0.5), 1)])
    cases["background"] = p
    p = tmp_path / "double.nvdb"
    _write_nvdb(p, [(_grid_bytes(root_tiles, bbox, MAT, VEC, grid_type=2), 2)])
    cases["no float grid"] = p
    for what, path in cases.items():
        with pytest.raises(mcpt.McptError):
            mcpt.load_nvdb(path)
    with pytest.raises(mcpt.McptError):
        mcpt.load_nvdb(tmp_path / "missing.nvdb")
