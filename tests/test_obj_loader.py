"""OBJ + MTL import (mcpt_load_obj, host code) against known answers of the
reference's rules (loadObject / loadMtlFile / parseObjFaceToken,
src/path_tracer.cu:2527-2859): fan triangulation, 1-based and negative
indices, the normalisation transform, area-weighted vertex normals, rotated
file normals, sRGB albedo, emissive Ke, the usemtl fallbacks and skipped
faces. Runs on the CPU (no GPU call)."""
import ctypes as C
import math

import numpy as np
import pytest

CUBE_V = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
CUBE_F = [(1, 2, 3, 4), (5, 8, 7, 6), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 8, 4), (5, 1, 4, 8)]


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _cube_obj(relative=False):
    lines = [f"v {x} {y} {z}" for x, y, z in CUBE_V]
    for f in CUBE_F:
        if relative:  # the same corners as negative indices (all 8 vertices read)
            lines.append("f " + " ".join(str(i - 9) for i in f))
        else:
            lines.append("f " + " ".join(str(i) for i in f))
    return "\n".join(lines) + "\n"


def _tri(hot):
    return np.stack([hot["v0"], hot["v1"], hot["v2"]], 1).astype(np.float64)


def test_fan_triangulation_and_normalisation(mcpt, tmp_path):
    hot, cold, mats, st = mcpt.load_obj(_write(tmp_path, "cube.obj", _cube_obj()))
    assert st["n_tris"] == 12 and st["skipped_faces"] == 0 and st["n_materials"] == 1
    # bottom centre (0.5, 0, 0.5) to the origin, largest side already 1
    v = np.array(CUBE_V, np.float64) - (0.5, 0.0, 0.5)
    want = []
    for f in CUBE_F:
        for k in range(1, len(f) - 1):  # fan (0, k, k + 1)
            want.append([v[f[0] - 1], v[f[k] - 1], v[f[k + 1] - 1]])
    assert np.array_equal(_tri(hot), np.array(want))
    assert np.all(cold["material"] == 0)
    # the relative-index form of the same file gives the same triangles
    hot2, _, _, _ = mcpt.load_obj(_write(tmp_path, "cube_rel.obj", _cube_obj(relative=True)))
    assert np.array_equal(hot, hot2)


def test_area_weighted_vertex_normals(mcpt, tmp_path):
    hot, cold, _, _ = mcpt.load_obj(_write(tmp_path, "cube.obj", _cube_obj()))
    t = _tri(hot)
    # sum of (v1 - v0) x (v2 - v0) over the triangles that use each vertex
    acc = {}
    for tri in t:
        a = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        for p in tri:
            acc[tuple(p)] = acc.get(tuple(p), 0) + a
    for k, key in enumerate(("n0", "n1", "n2")):
        for i in range(len(t)):
            n = acc[tuple(t[i, k])]
            assert np.allclose(cold[key][i], n / np.linalg.norm(n), atol=2e-7)
    # unit length, and every corner normal on the same side (the faces'
    # winding) of the line from the cube's centre through that corner
    assert np.allclose(np.linalg.norm(cold["n0"], axis=1), 1.0, atol=1e-6)
    side = np.sign(np.einsum("ij,ij->i", cold["n0"], t[:, 0] - np.array([0.0, 0.5, 0.0])))
    assert np.all(side == side[0]) and side[0] != 0


def test_transform_and_rotated_normals(mcpt, tmp_path):
    text = "\n".join([
        "v 0 0 0", "v 4 0 0", "v 4 2 0", "v 0 2 0",
        "vn 0 0 2", "vt 0.25 0.75", "vt 1 0",
        "f 1/1/1 2/2/1 3//1", "f 1//1 3//1 4//1",
    ]) + "\n"
    ang = 0.7
    hot, cold, _, st = mcpt.load_obj(_write(tmp_path, "q.obj", text), translation=(1, 2, 3), scale=2.0,
                                     rotation_y=ang)
    assert st["n_tris"] == 2
    t = _tri(hot).reshape(-1, 3)
    # largest side 4 -> uniform_scale 2; bottom at y = 2; rotation keeps the
    # x-z distances from the bottom centre (1, 2, 3)
    assert t[:, 1].min() == pytest.approx(2.0, abs=1e-6)
    assert t[:, 1].max() == pytest.approx(3.0, abs=1e-6)
    # x in [0, 4] around the bottom centre x = 2, times 2/4: one unit away
    r = np.hypot(t[:, 0] - 1, t[:, 2] - 3)
    assert np.allclose(r, 1.0, atol=1e-6)
    # the file normal (0,0,2) rotated about +y by ang and normalised
    n = np.array([math.sin(ang), 0.0, math.cos(ang)])
    assert np.allclose(cold["n0"], n, atol=1e-6) and np.allclose(cold["n2"], n, atol=1e-6)
    # texture coordinates where given, zero where not
    assert np.allclose(cold["uv0"][0], (0.25, 0.75)) and np.allclose(cold["uv1"][0], (1, 0))
    assert np.all(cold["uv2"][0] == 0)


def test_polygon_fan_and_skipped_faces(mcpt, tmp_path):
    text = "\n".join([
        "v 0 0 0", "v 1 0 0", "v 1.5 1 0", "v 0.5 1.6 0", "v -0.5 1 0",
        "v 2 0 0", "v 3 0 0",
        "f 1 2 3 4 5",       # pentagon: 3 triangles
        "f 1 2 6",           # collinear: zero area, skipped
        "f 1 2 9",           # index beyond the list: skipped
        "f 1 2",             # too few corners: no triangle
    ]) + "\n"
    hot, _, _, st = mcpt.load_obj(_write(tmp_path, "p.obj", text))
    assert st["n_tris"] == 3 and st["skipped_faces"] == 2


def test_mtl_materials(mcpt, tmp_path):
    _write(tmp_path, "m.mtl", "\n".join([
        "newmtl grey", "Kd 0.5 0.5 0.5",
        "newmtl lamp", "Kd 0.2 0.2 0.2", "Ke 4 5 6",
        "newmtl tex", "Kd 0.04 0.5 1", "map_Kd textures\\wall.png",
    ]) + "\n")
    text = "\n".join([
        "mtllib m.mtl missing.mtl",
        "v 0 0 0", "v 1 0 0", "v 0 1 0", "v 0 0 1",
        "usemtl lamp", "f 1 2 3",
        "usemtl tex", "f 1 2 4",
        "usemtl nowhere", "f 1 3 4",   # unknown, library of 3: fallback
        "usemtl grey", "f 2 3 4",
    ]) + "\n"
    hot, cold, mats, st = mcpt.load_obj(_write(tmp_path, "m.obj", text))
    assert st["n_materials"] == 4 and st["texture_requests"] == 1 and st["missing_libraries"] == 1
    assert list(cold["material"]) == [2, 3, 0, 1]

    def lin(e):  # srgbToLinear, closed form
        return e / 12.92 if e <= 0.04045 else ((e + 0.055) / 1.055) ** 2.4

    assert mats[1]["type"] == mcpt.MAT_LAMBERTIAN
    assert np.allclose(mats[1]["albedo"], lin(0.5), rtol=1e-6) and abs(lin(0.5) - 0.2140411) < 1e-6
    assert mats[2]["type"] == mcpt.MAT_EMISSIVE and np.array_equal(mats[2]["emission"], [4, 5, 6])
    assert np.all(mats[2]["albedo"] == 0)
    assert mats[3]["type"] == mcpt.MAT_LAMBERTIAN  # the texture is not loaded: Kd stays
    assert np.allclose(mats[3]["albedo"], [0.04 / 12.92, lin(0.5), 1.0], rtol=1e-6)
    assert mats[0]["type"] == mcpt.MAT_LAMBERTIAN and np.allclose(mats[0]["albedo"], 0.73)


def test_unknown_material_with_single_library_entry(mcpt, tmp_path):
    _write(tmp_path, "one.mtl", "newmtl only\nKd 1 0 0\n")
    text = "mtllib one.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl other\nf 1 2 3\n"
    _, cold, mats, _ = mcpt.load_obj(_write(tmp_path, "o.obj", text))
    assert cold["material"][0] == 1 and np.allclose(mats[1]["albedo"], [1, 0, 0])


def test_errors(mcpt, tmp_path):
    with pytest.raises(mcpt.McptError):
        mcpt.load_obj(str(tmp_path / "absent.obj"))
    with pytest.raises(mcpt.McptError):
        mcpt.load_obj(_write(tmp_path, "empty.obj", "# nothing\nv 0 0 0\n"))
    with pytest.raises(mcpt.McptError):
        mcpt.load_obj(_write(tmp_path, "cube.obj", _cube_obj()), scale=0.0)


def test_round_trip_generated_mesh(mcpt, tmp_path):
    """A generated mesh written with 9 significant digits (exact for
    float32) loads back as the reference's float32 normalisation of the
    same vertices: (p - bottom centre) * (1 / largest side), bit for bit."""
    hot, _, _ = mcpt.gen_scene(mcpt.GEN_SOUP, seed=4, param=20000)
    v = np.stack([hot["v0"], hot["v1"], hot["v2"]], 1).reshape(-1, 3)
    lines = ["v %.9g %.9g %.9g" % tuple(p) for p in v]
    lines += ["f %d %d %d" % (3 * i + 1, 3 * i + 2, 3 * i + 3) for i in range(len(hot))]
    got, _, _, st = mcpt.load_obj(_write(tmp_path, "soup.obj", "\n".join(lines) + "\n"))
    assert st["n_tris"] == len(hot) and st["skipped_faces"] == 0
    lo, hi = v.min(0), v.max(0)
    f32 = np.float32
    bottom = np.array([(lo[0] + hi[0]) * f32(0.5), lo[1], (lo[2] + hi[2]) * f32(0.5)], f32)
    scale = f32(1.0) / (hi - lo).max()
    want = ((v - bottom) * scale).reshape(-1, 3, 3)
    assert np.array_equal(np.stack([got["v0"], got["v1"], got["v2"]], 1), want)


def _ppm(path, img):
    h, w = img.shape[:2]
    path.write_bytes(b"P6\n%d %d\n255\n" % (w, h) + np.ascontiguousarray(img[..., :3]).tobytes())


def test_map_kd_textures(mcpt, tmp_path):
    """map_Kd with load_textures (convertTexturePathToPpm and
    getOrLoadObjTexture, src/path_tracer.cu:2580-2611, 2636-2646): the path's
    backslashes become '/', its extension becomes .ppm, it is read relative
    to the OBJ's folder once per path, a material whose image loads gets
    that image, one whose image is missing keeps Kd."""
    (tmp_path / "tex").mkdir()
    rng = np.random.default_rng(3)
    a = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    b = rng.integers(0, 256, (4, 2, 3), dtype=np.uint8)
    _ppm(tmp_path / "tex" / "wall.ppm", a)
    _ppm(tmp_path / "tex" / "b.v2.ppm", b)
    _write(tmp_path, "m.mtl", "\n".join([
        "newmtl w1", "Kd 0.5 0.5 0.5", "map_Kd tex\\wall.png",
        "newmtl w2", "Kd 0.1 0.1 0.1", "map_Kd tex/wall.jpg",      # same .ppm: shared
        "newmtl gone", "Kd 0.2 0.3 0.4", "map_Kd tex/none.png",    # missing: keeps Kd
        "newmtl lamp", "Ke 1 1 1", "map_Kd tex/wall.png",          # emissive wins
        "newmtl dotted", "Kd 1 1 1", "map_Kd tex/b.v2.tga",
    ]) + "\n")
    obj = "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\n" + "".join(
        f"usemtl {m}\nf 1 2 3\n" for m in ("w1", "w2", "gone", "lamp", "dotted"))
    path = _write(tmp_path, "t.obj", obj)
    hot, cold, mats, st, images = mcpt.load_obj(path, load_textures=True)
    assert st["texture_requests"] == 4 and st["textures_loaded"] == 2 and st["missing_textures"] == 1
    assert len(images) == 2
    for img, want in zip(images, (a, b)):
        assert np.array_equal(img[..., :3], want) and np.all(img[..., 3] == 255)
    assert list(mats["texture"][1:]) == [1, 1, 0, 0, 2]
    assert np.allclose(mats[3]["albedo"], [0.0331048, 0.0732390, 0.1328683], rtol=1e-5)  # sRGB 0.2, 0.3, 0.4
    # without load_textures nothing is read and every material keeps Kd
    _, _, mats2, st2 = mcpt.load_obj(path)
    assert np.all(mats2["texture"] == 0) and st2["textures_loaded"] == 0 and st2["texture_requests"] == 4
    # the images handed out are those of the thread's last load, which has none
    w, h = C.c_int32(), C.c_int32()
    assert mcpt._lib.mcpt_obj_texture(0, C.byref(w), C.byref(h), None) != 0
