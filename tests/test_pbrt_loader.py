"""pbrt-v3 import (mcpt_load_pbrt, host code) against known answers of the
pbrt-v3 file format and of the reference's mapping (loadPbrtScene and its
helpers, src/path_tracer.cu:1990-2483). The reference parses with minipbrt,
which is not in this image; every expectation below is a closed form of the
format or of the reference's mapping code. Runs on the CPU (no GPU call)."""
import math
import struct

import numpy as np
import pytest


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _tris(hot):
    return np.stack([hot["v0"], hot["v1"], hot["v2"]], 1).astype(np.float64)


def test_camera_film_and_transforms(mcpt, tmp_path):
    scene = """
    LookAt 0 0 5  0 0 0  0 1 0   # camera at +5 z looking at the origin
    Camera "perspective" "float fov" [ 60 ]
    Film "image" "integer xresolution" [ 640 ] "integer yresolution" [ 480 ]
    WorldBegin
    AttributeBegin
      Material "matte" "rgb Kd" [ 0.2 0.4 0.6 ]
      Translate 1 2 3
      Rotate 90 0 0 1
      Scale 2 2 2
      Shape "trianglemesh" "integer indices" [0 1 2 0 2 3]
            "point P" [0 0 0  1 0 0  1 1 0  0 1 0] "float uv" [0 0 1 0 1 1 0 1]
            "normal N" [0 0 1  0 0 1  0 0 1  0 0 1]
    AttributeEnd
    TransformBegin
      ConcatTransform [1 0 0 0  0 1 0 0  0 0 1 0  10 20 30 1]
      Shape "trianglemesh" "point P" [0 0 0  1 0 0  0 1 0]
    TransformEnd
    Shape "trianglemesh" "point P" [0 0 0  1 0 0  0 1 0]
    WorldEnd
    """
    hot, cold, sph, mats, info, _ = mcpt.load_pbrt(_write(tmp_path, "a.pbrt", scene))
    assert info["has_camera"] and info["width"] == 640 and info["height"] == 480
    assert np.allclose(info["look_from"], [0, 0, 5], atol=1e-6)
    assert np.allclose(info["look_at"], [0, 0, 4], atol=1e-6)  # look_from + the camera's +z
    assert np.allclose(info["up"], [0, 1, 0], atol=1e-6) and info["vfov_degrees"] == pytest.approx(60)
    assert len(hot) == 4 and len(sph) == 0
    # T(1,2,3) R(90 about z) S(2): (x, y) -> (1 - 2y, 2 + 2x)
    t = _tris(hot)
    quad = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64)
    want = np.stack([1 - 2 * quad[:, 1], 2 + 2 * quad[:, 0], np.full(4, 3.0)], 1)
    assert np.allclose(t[0], want[[0, 1, 2]], atol=1e-5) and np.allclose(t[1], want[[0, 2, 3]], atol=1e-5)
    assert np.allclose(cold["n0"][0], [0, 0, 1]) and np.allclose(cold["uv2"][1], [0, 1])
    assert np.allclose(t[2], [[10, 20, 30], [11, 20, 30], [10, 21, 30]])  # column-major ConcatTransform
    assert np.allclose(t[3], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])           # TransformEnd restored identity
    assert np.allclose(mats[cold["material"][0]]["albedo"], [0.2, 0.4, 0.6])  # linear, no sRGB conversion
    # a portrait film: pbrt's fov belongs to the shorter side
    hot, _, _, _, info, _ = mcpt.load_pbrt(_write(tmp_path, "b.pbrt", scene.replace("640", "300")))
    want = 2 * math.degrees(math.atan(math.tan(math.radians(30)) * 480 / 300))
    assert info["vfov_degrees"] == pytest.approx(want, rel=1e-5)


def test_materials(mcpt, tmp_path):
    tri = 'Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0]\n'
    scene = "WorldBegin\n" + "".join([
        'Material "matte"\n', tri,                                   # Kd 0.5
        'Material "plastic"\n', tri,                                 # Kd 0.25
        'Material "glass" "float eta" 1.33\n', tri,
        'Material "mirror" "rgb Kr" [0.7 0.8 0.9]\n', tri,
        'Material "metal" "rgb eta" [0.2 0.9 1.1] "rgb k" [3.9 2.4 2.1] "float roughness" 0.3\n', tri,
        'MakeNamedMaterial "red" "string type" "matte" "rgb Kd" [0.9 0.1 0.1]\n',
        'MakeNamedMaterial "mixed" "string type" "mix" "string namedmaterial1" "red" '
        '"string namedmaterial2" "red"\n',
        'NamedMaterial "mixed"\n', tri,
        'Material "none"\n', tri,                                    # invisible interface: skipped
        'NamedMaterial "unknown"\n', tri,                            # grey 0.5 fallback
    ]) + "WorldEnd\n"
    hot, cold, _, mats, info, _ = mcpt.load_pbrt(_write(tmp_path, "m.pbrt", scene))
    assert len(hot) == 7 and info["skipped_shapes"] == 1
    got = [mats[k] for k in cold["material"]]
    assert got[0]["type"] == mcpt.MAT_LAMBERTIAN and np.allclose(got[0]["albedo"], 0.5)
    assert np.allclose(got[1]["albedo"], 0.25)
    assert got[2]["type"] == mcpt.MAT_DIELECTRIC and got[2]["ior"] == pytest.approx(1.33)
    assert got[3]["type"] == mcpt.MAT_CONDUCTOR and np.allclose(got[3]["albedo"], [0.7, 0.8, 0.9])
    eta, k = np.array([0.2, 0.9, 1.1]), np.array([3.9, 2.4, 2.1])
    assert got[4]["type"] == mcpt.MAT_CONDUCTOR and got[4]["fuzz"] == pytest.approx(0.3)
    assert np.allclose(got[4]["albedo"], ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k), rtol=1e-5)
    assert np.allclose(got[5]["albedo"], [0.9, 0.1, 0.1])  # mix takes its first material
    assert np.allclose(got[6]["albedo"], 0.5)
    assert len(mats) == 6  # identical records share one entry (matte 0.5 and the fallback)


def test_lights_media_textures(mcpt, tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (3, 4, 3), dtype=np.uint8)
    (tmp_path / "wood.ppm").write_bytes(b"P6\n4 3\n255\n" + img.tobytes())
    tri = 'Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0]\n'
    scene = "WorldBegin\n" + "".join([
        'LightSource "infinite" "rgb L" [0.1 0.2 0.3]\n',
        'AttributeBegin\n', 'AreaLightSource "diffuse" "rgb L" [4 5 6] "rgb scale" [2 2 2]\n', tri, 'AttributeEnd\n',
        'MakeNamedMedium "fog" "string type" "homogeneous" "rgb sigma_a" [0.1 0.2 0.3] '
        '"rgb sigma_s" [0.5 0.5 0.5] "float scale" 2 "float g" 0.4\n',
        'AttributeBegin\n', 'MediumInterface "fog" ""\n', tri, 'AttributeEnd\n',
        'Texture "w" "spectrum" "imagemap" "string filename" "wood.ppm"\n',
        'Texture "gone" "spectrum" "imagemap" "string filename" "missing.png"\n',
        'Texture "c" "spectrum" "constant" "rgb value" [0.3 0.6 0.9]\n',
        'Texture "s" "spectrum" "scale" "texture tex1" "c" "rgb tex2" [2 2 2]\n',
        'Material "matte" "texture Kd" "w"\n', tri,
        'Material "matte" "texture Kd" "gone"\n', tri,
        'Material "matte" "texture Kd" "s"\n', tri,
        'LightSource "point" "point from" [0 5 0] "rgb I" [1 1 1]\n',
        'LightSource "distant" "point from" [0 1 0] "point to" [0 0 0] "rgb L" [3 3 3]\n',
    ]) + "WorldEnd\n"
    hot, cold, _, mats, info, images = mcpt.load_pbrt(_write(tmp_path, "l.pbrt", scene))
    assert np.allclose(info["background"], [0.1, 0.2, 0.3]) and not info["has_env_map"]
    got = [mats[k] for k in cold["material"]]
    assert got[0]["type"] == mcpt.MAT_EMISSIVE and np.allclose(got[0]["emission"], [8, 10, 12])
    # pbrtMediumMaterial: albedo sigma_s / sigma_t, density mean sigma_t
    sa, ss = np.array([0.1, 0.2, 0.3]) * 2, np.array([0.5, 0.5, 0.5]) * 2
    assert got[1]["type"] == mcpt.MAT_ISOTROPIC and got[1]["phase_g"] == pytest.approx(0.4)
    assert np.allclose(got[1]["albedo"], ss / (sa + ss), rtol=1e-6)
    assert got[1]["density"] == pytest.approx(float(np.mean(sa + ss)), rel=1e-6)
    assert got[2]["texture"] == 1 and len(images) == 1 and np.array_equal(images[0][..., :3], img)
    assert got[3]["texture"] == 0 and np.allclose(got[3]["albedo"], 0.5) and info["unsupported_textures"] == 1
    assert np.allclose(got[4]["albedo"], [0.3, 0.6, 0.9])  # scale: its first texture
    # the point light: a box of 12 emissive triangles, half size 0.5% of the
    # scene's half diagonal, radiance I / (6 h^2); the distant light one quad
    assert len(hot) == 5 + 12 + 2 and info["approximated_lights"] == 2
    box = _tris(hot[5:17]).reshape(-1, 3)
    h = (box.max(0) - box.min(0))[0] / 2
    lo, hi = _tris(hot[:5]).reshape(-1, 3).min(0), _tris(hot[:5]).reshape(-1, 3).max(0)
    assert h == pytest.approx(np.linalg.norm(hi - lo) / 2 * 0.005, rel=1e-5)
    assert np.allclose(box.mean(0), [0, 5, 0], atol=1e-5)
    assert mats[cold["material"][5]]["emission"][0] == pytest.approx(1 / (6 * h * h), rel=1e-4)
    t = math.tan(math.radians(1))
    assert mats[cold["material"][17]]["emission"][0] == pytest.approx(3 / (4 * t * t), rel=1e-4)
    # an infinite light with a map gives the path and the rotation
    hot, _, _, _, info, _ = mcpt.load_pbrt(_write(tmp_path, "e.pbrt", "WorldBegin\n"
                                                  'LightSource "infinite" "string mapname" "sky.pfm"\n' + tri))
    assert info["has_env_map"] and info["env_path"].endswith("sky.pfm")
    assert info["env_rotation_degrees"] == pytest.approx(-90.0)  # -90 deg - atan2(0, 0)
    # the light's +y turned to world +z: -90 - atan2(1, 0) = -180 degrees
    _, _, _, _, info, _ = mcpt.load_pbrt(_write(tmp_path, "f.pbrt", "WorldBegin\nAttributeBegin\nRotate 90 1 0 0\n"
                                              'LightSource "infinite" "string mapname" "sky.pfm"\nAttributeEnd\n' + tri))
    assert info["env_rotation_degrees"] == pytest.approx(-180.0, abs=1e-4)


def test_spheres_instances_include(mcpt, tmp_path):
    _write(tmp_path, "part.pbrt", 'Shape "sphere" "float radius" 2\n')
    scene = """WorldBegin
    ObjectBegin "tri"
      Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0]
    ObjectEnd
    AttributeBegin Translate 5 0 0 ObjectInstance "tri" AttributeEnd
    AttributeBegin Translate 0 7 0 ObjectInstance "tri" AttributeEnd
    AttributeBegin Translate 1 1 1 Scale 1 2 3 Include "part.pbrt" AttributeEnd
    AttributeBegin AreaLightSource "diffuse" "rgb L" [1 1 1] Shape "sphere" "float radius" 0.5 AttributeEnd
    Shape "cone"
    WorldEnd"""
    hot, cold, sph, mats, info, _ = mcpt.load_pbrt(_write(tmp_path, "s.pbrt", scene))
    assert info["instances"] == 2 and info["skipped_shapes"] == 1
    # shapes first, instances after them (the reference's order, :2367-2386)
    t = _tris(hot)
    assert np.allclose(t[-2], [[5, 0, 0], [6, 0, 0], [5, 1, 0]]) and np.allclose(t[-1], [[0, 7, 0], [1, 7, 0], [0, 8, 0]])
    # the included sphere: centre at the translation, radius scaled by the mean axis length
    assert len(sph) == 1 and np.allclose(sph["center"][0], [1, 1, 1]) and sph["radius"][0] == pytest.approx(4.0)
    # the emissive sphere: tessellated on 16 parallels x 32 meridians, minus
    # the degenerate triangles at the top pole
    lamp = _tris(hot[:-2]).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(lamp, axis=1), 0.5, rtol=1e-5)
    assert 32 * 16 * 2 - 64 <= len(hot) - 2 <= 32 * 16 * 2 - 32
    assert all(mats[k]["type"] == mcpt.MAT_EMISSIVE for k in cold["material"][:-2])


def _ply_binary(verts, faces, normals=None, uv=None):
    props = ["x", "y", "z"] + (["nx", "ny", "nz"] if normals is not None else []) + (["u", "v"] if uv is not None else [])
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(verts)}"]
    head += [f"property float {p}" for p in props]
    head += [f"element face {len(faces)}", "property list uchar int vertex_indices", "end_header"]
    body = b""
    for i, v in enumerate(verts):
        row = list(v) + (list(normals[i]) if normals is not None else []) + (list(uv[i]) if uv is not None else [])
        body += struct.pack("<%df" % len(row), *row)
    for f in faces:
        body += struct.pack("<B%di" % len(f), len(f), *f)
    return ("\n".join(head) + "\n").encode() + body


def test_ply_meshes(mcpt, tmp_path):
    verts = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    normals = [(0, 0, 2)] * 4
    uv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    (tmp_path / "q.ply").write_bytes(_ply_binary(verts, [(0, 1, 2, 3)], normals, uv))
    ascii_ply = "\n".join(["ply", "format ascii 1.0", "comment a test", "element vertex 3", "property float x",
                           "property float y", "property float z", "element face 1",
                           "property list uchar int vertex_indices", "end_header", "0 0 0", "2 0 0", "0 2 0",
                           "3 0 1 2"]) + "\n"
    _write(tmp_path, "t.ply", ascii_ply)
    scene = ('WorldBegin\nTranslate 0 0 1\nShape "plymesh" "string filename" "q.ply"\n'
             'Shape "plymesh" "string filename" "t.ply"\nWorldEnd\n')
    hot, cold, _, _, _, _ = mcpt.load_pbrt(_write(tmp_path, "p.pbrt", scene))
    t = _tris(hot)
    assert len(hot) == 3  # the quad fanned into two, the triangle
    assert np.allclose(t[0], [[0, 0, 1], [1, 0, 1], [1, 1, 1]]) and np.allclose(t[1], [[0, 0, 1], [1, 1, 1], [0, 1, 1]])
    assert np.allclose(cold["n0"][:2], [0, 0, 1]) and np.allclose(cold["uv1"][0], [1, 0])
    assert np.allclose(t[2], [[0, 0, 1], [2, 0, 1], [0, 2, 1]]) and np.all(cold["n0"][2] == 0)


def test_errors(mcpt, tmp_path):
    with pytest.raises(mcpt.McptError):
        mcpt.load_pbrt(str(tmp_path / "absent.pbrt"))
    for bad in ('WorldBegin\nShape "trianglemesh" "point P" [0 0 0 1 0 0\n',
                'WorldBegin\nShape "trianglemesh\n',
                'WorldBegin\nAttributeEnd\n',
                'WorldBegin\nShape "cone"\nWorldEnd\n',  # nothing renderable
                'WorldBegin\nShape "plymesh" "string filename" "none.ply"\n'):
        with pytest.raises(mcpt.McptError):
            mcpt.load_pbrt(_write(tmp_path, "bad.pbrt", bad))
