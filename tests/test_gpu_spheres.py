"""Spheres, motion blur and the thin lens on the GPU against the CPU oracle.

The build over finalizeScene's primitive list (spheres first, then the
triangles, src/path_tracer.cu:1955-1972) is bit-exact stage by stage; closest
and any hits (hit_sphere at the ray's time, :886-911) are bit-exact in t,
primitive type and index in both traversal orders; renders of a room holding
Lambertian (image-textured), metal, glass, emissive, smoke (constant medium)
and moving spheres are bit-identical across the render modes and within
relMSE 1e-4 of the oracle, and so is a render through a thin lens
(Camera::getRay's lens disk, :1648-1658).
"""
import numpy as np
import pytest

from conftest import rel_mse
from test_gpu_parity import LBVH_KEYS, _bounds_equal, _render_modes

pytestmark = pytest.mark.gpu


def _random_scene(mcpt, rng, n_sph, n_tri, moving_every=3):
    sp = np.zeros(n_sph, mcpt.SPHERE_DTYPE)
    sp["center"] = rng.uniform(-4, 4, (n_sph, 3))
    sp["radius"] = rng.uniform(0.05, 0.5, n_sph)
    sp["moving"][::moving_every] = 1
    sp["center1"] = sp["center"] + rng.uniform(-1, 1, (n_sph, 3)) * sp["moving"][:, None]
    if n_tri:
        hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=n_tri, param=n_tri)
        for k in ("v0", "v1", "v2"):
            hot[k] = hot[k] * np.float32(8) - np.float32(4)
    else:
        hot, cold = np.zeros(0, mcpt.HOT_DTYPE), np.zeros(0, mcpt.COLD_DTYPE)
        mats = np.zeros(1, mcpt.MAT_DTYPE)
        mats["albedo"] = 0.5
    return hot, cold, mats, sp


def _rays(rng, n):
    rays = np.zeros(n, np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")]))
    o = rng.uniform(-6, 6, (n, 3))
    d = rng.uniform(-3, 3, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["origin"], rays["dir"], rays["tmax"] = o, d, 1e30
    rays["time"] = rng.uniform(0, 1, n)
    return rays


@pytest.mark.parametrize("n_sph,n_tri", [(1, 0), (2, 0), (300, 0), (257, 4096), (5000, 65536)])
def test_build_with_spheres(mcpt, ctx, oracle, n_sph, n_tri):
    rng = np.random.default_rng(n_sph + n_tri)
    hot, cold, mats, sp = _random_scene(mcpt, rng, n_sph, n_tri)
    print(f"sphere build parity spheres={n_sph} tris={len(hot)}", flush=True)
    ref = oracle.Scene(hot, cold, mats, spheres=sp).dump()
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    got = scene.build(dump=True)
    for k in LBVH_KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert _bounds_equal(got["internal_bounds"], ref["internal_bounds"])
    assert _bounds_equal(got["leaf_bounds"], ref["leaf_bounds"])
    assert _bounds_equal(got["scene_bounds"], ref["scene_bounds"])
    assert got["max_depth"] == ref["max_depth"]
    scene.close()


@pytest.mark.parametrize("n_sph,n_tri", [(300, 0), (2000, 16384)])
def test_trace_with_spheres(mcpt, ctx, oracle, n_sph, n_tri):
    rng = np.random.default_rng(7 * n_sph + n_tri)
    hot, cold, mats, sp = _random_scene(mcpt, rng, n_sph, n_tri)
    print(f"sphere trace parity spheres={n_sph} tris={len(hot)}", flush=True)
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    scene.build()
    ref = oracle.Scene(hot, cold, mats, spheres=sp)
    rays = _rays(rng, 1 << 15)
    hits, steps, leaves = scene.trace_closest(rays, counts=True)
    rh, rs, rl = ref.trace_closest(rays)
    for k in ("prim_type", "prim_index", "did_hit"):
        assert np.array_equal(hits[k], rh[k]), k
    assert np.array_equal(hits["t"].view(np.uint32), rh["t"].view(np.uint32))
    assert np.array_equal(steps, rs) and np.array_equal(leaves, rl)
    assert np.array_equal(scene.trace_closest(rays), hits)  # the four-wide walk
    assert (hits["prim_type"] == 0).sum() > 1000
    rays["tmax"] = np.float32(1.5)
    occ = scene.trace_any(rays)
    assert np.array_equal(occ, ref.trace_any(rays)) and 0 < occ.mean() < 1
    scene.close()


def _sphere_room(mcpt, rng):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    extra = np.zeros(6, mcpt.MAT_DTYPE)
    extra[0]["type"], extra[0]["texture"] = mcpt.MAT_LAMBERTIAN, 1               # earth-like, image texture
    extra[1]["type"], extra[1]["albedo"], extra[1]["fuzz"] = mcpt.MAT_CONDUCTOR, (0.8, 0.8, 0.9), 0.05
    extra[2]["type"], extra[2]["ior"] = mcpt.MAT_DIELECTRIC, 1.5
    extra[3]["type"], extra[3]["emission"] = mcpt.MAT_EMISSIVE, (4.0, 2.0, 1.0)
    extra[4]["type"], extra[4]["albedo"], extra[4]["density"], extra[4]["phase_g"] = mcpt.MAT_ISOTROPIC, 0.9, 0.8, 0.3
    extra[5]["type"], extra[5]["albedo"] = mcpt.MAT_LAMBERTIAN, (0.2, 0.6, 0.3)
    m0 = len(mats)
    mats = np.concatenate([mats, extra])
    sp = np.zeros(6, mcpt.SPHERE_DTYPE)
    sp["center"] = [(-1.0, 0.7, -2.5), (1.0, 0.8, -2.8), (0.2, 0.5, -1.2), (-1.2, 2.8, -3.0), (0.9, 2.0, -1.8),
                    (-0.6, 0.4, -0.9)]
    sp["radius"] = [0.7, 0.8, 0.5, 0.3, 0.6, 0.4]
    sp["material"] = m0 + np.arange(6)
    sp["moving"][5] = 1
    sp["center1"][5] = (0.4, 0.6, -0.9)
    img = rng.integers(0, 256, (32, 64, 4), dtype=np.uint8)
    return hot, cold, mats, sp, img


@pytest.mark.parametrize("aperture", [0.0, 0.4])
def test_render_sphere_room(mcpt, ctx, oracle, aperture):
    rng = np.random.default_rng(31)
    hot, cold, mats, sp, img = _sphere_room(mcpt, rng)
    W = H = 64
    spp, depth, seed = 4, 8, 17
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    desc.aperture = aperture
    desc.focus_distance = 7.0
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    ref = oracle.Scene(hot, cold, mats, spheres=sp)
    for target in (scene, ref):
        assert target.add_texture(mcpt.TEX_IMAGE, image=img) == 0
    scene.build()
    print(f"sphere room render aperture={aperture}", flush=True)
    ref_img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    assert ref_img.mean() > 0.01
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (im, _) in out.items():
        assert np.array_equal(im, base), mode
        assert rel_mse(im, ref_img) < 1e-4, mode
        print(f"sphere room mode {mode}: identical to the oracle {np.mean(im == ref_img):.5f}", flush=True)
    scene.close()


def _alone(mcpt, ctx, scene, cam, spp, depth, seed):
    film = ctx.film(cam.width, cam.height)
    scene.render(cam, film, spp=spp, max_depth=depth, seed=seed)
    img = film.read()
    film.close()
    return img


def test_batch_and_server_with_spheres(mcpt, ctx):
    """The batched build (sphere scenes take the single-scene build), the
    batched render and the server give each sphere scene, moving spheres
    included, the film mcpt_render gives it alone, bit for bit: their shared
    queues carry each ray's scene and time together."""
    rng = np.random.default_rng(12)
    hot, cold, mats, sp, _ = _sphere_room(mcpt, rng)
    mats["texture"] = 0
    h2, c2, m2, sp2 = _random_scene(mcpt, rng, 200, 4096)
    h3, c3, m3, sp3 = _random_scene(mcpt, rng, 50, 1024)
    sp3["moving"] = 0
    soup = mcpt.gen_scene(mcpt.GEN_SOUP, seed=3, param=3000)
    scenes = [ctx.scene(hot, cold, mats, spheres=sp), ctx.scene(h2, c2, m2, spheres=sp2),
              ctx.scene(h3, c3, m3, spheres=sp3), ctx.scene(*soup)]
    mcpt.build_batch(scenes)
    cams = [mcpt.camera_init(mcpt.gen_camera(k, w, h)) for k, w, h in
            ((mcpt.GEN_CORNELL, 48, 40), (mcpt.GEN_SOUP, 32, 32), (mcpt.GEN_SOUP, 20, 12), (mcpt.GEN_SOUP, 24, 16))]
    spp, depth, seed = 3, 8, 4
    refs = [_alone(mcpt, ctx, sc, cam, spp, depth, seed) for sc, cam in zip(scenes, cams)]
    films = [ctx.film(c.width, c.height) for c in cams]
    mcpt.render_batch(scenes, cams, films, spp=spp, max_depth=depth, seed=seed)
    for i, (f, ref) in enumerate(zip(films, refs)):
        assert np.array_equal(f.read(), ref), i
        f.close()
    print("batched render with spheres: films identical", flush=True)
    # the static-sphere and triangle scenes alone use the scene-only records
    films = [ctx.film(c.width, c.height) for c in cams[2:]]
    mcpt.render_batch(scenes[2:], cams[2:], films, spp=spp, max_depth=depth, seed=seed)
    for i, (f, ref) in enumerate(zip(films, refs[2:])):
        assert np.array_equal(f.read(), ref), i
        f.close()
    srv = ctx.server(scenes)
    ids = [srv.submit(i, cam, spp, max_depth=depth, seed=seed)[0] for i, cam in enumerate(cams)]
    for i, jid in enumerate(ids):
        res, img = srv.wait(jid)
        assert res["status"] == mcpt.OK and np.array_equal(img, refs[i]), i
        srv.release(jid)
    srv.close()
    print("server with spheres: films identical", flush=True)
    for sc in scenes:
        sc.close()


@pytest.mark.parametrize("frame_spp", [1, 3])
def test_cache_with_spheres(mcpt, ctx, frame_spp):
    """The radiance cache on a room with spheres (moving ones included):
    its first, untrained frame is the plain render bit for bit, and after
    training it queries the network."""
    rng = np.random.default_rng(31)
    hot, cold, mats, sp, img = _sphere_room(mcpt, rng)
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    assert scene.add_texture(mcpt.TEX_IMAGE, image=img) == 0
    scene.build()
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 40, 32))
    net = ctx.mlp(seed=2)
    cached, plain = ctx.film(40, 32), ctx.film(40, 32)
    st = scene.render_cached(cam, cached, net, spp=frame_spp, sample_begin=5, max_depth=8, seed=7, train_paths=256,
                             train_batch=512, frame_spp=frame_spp)
    assert st["queries"] == 0 and st["train_samples"] > 0
    scene.render(cam, plain, mode=mcpt.MODE_WAVEFRONT, spp=frame_spp, sample_begin=5, max_depth=8, seed=7)
    assert np.array_equal(cached.read(), plain.read())
    st = scene.render_cached(cam, cached, net, spp=2, sample_begin=20, max_depth=8, seed=7, train_paths=256,
                             train_batch=512)
    assert st["queries"] > 0 and np.all(np.isfinite(cached.read()))
    for f in (cached, plain):
        f.close()
    net.close()
    scene.close()


PBRT_ROOM = """
LookAt 0 1 3.2  0 1 0  0 1 0
Camera "perspective" "float fov" [ 45 ]
Film "image" "integer xresolution" [ 64 ] "integer yresolution" [ 48 ]
WorldBegin
Material "matte" "rgb Kd" [ 0.7 0.7 0.7 ]
Shape "trianglemesh" "integer indices" [0 1 2 0 2 3  4 5 6 4 6 7  0 4 7 0 7 3]
      "point P" [-1 0 -1  1 0 -1  1 0 1  -1 0 1   -1 2 -1  1 2 -1  1 2 1  -1 2 1]
AttributeBegin
  Material "matte" "rgb Kd" [ 0.6 0.1 0.1 ]
  Shape "trianglemesh" "integer indices" [0 1 2 0 2 3] "point P" [-1 0 -1  -1 2 -1  -1 2 1  -1 0 1]
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 12 12 12 ]
  Shape "trianglemesh" "integer indices" [0 1 2 0 2 3] "point P" [-0.3 1.98 -0.3  0.3 1.98 -0.3  0.3 1.98 0.3  -0.3 1.98 0.3]
AttributeEnd
AttributeBegin
  Material "glass" "float eta" 1.5
  Translate -0.4 0.35 0.1
  Shape "sphere" "float radius" 0.35
AttributeEnd
AttributeBegin
  Material "metal" "float roughness" 0.05
  Translate 0.45 0.4 -0.3
  Shape "sphere" "float radius" 0.4
AttributeEnd
WorldEnd
"""


def test_render_pbrt_scene(mcpt, ctx, oracle, tmp_path):
    """A pbrt-v3 room read by mcpt_load_pbrt (triangles, an area light,
    glass and metal spheres, the file's camera and film) renders the same
    on the GPU in every mode and in the oracle."""
    p = tmp_path / "room.pbrt"
    p.write_text(PBRT_ROOM)
    hot, cold, sp, mats, info, _ = mcpt.load_pbrt(str(p))
    assert info["has_camera"] and len(sp) == 2 and len(hot) == 10
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, info["width"], info["height"])
    desc.look_from[:], desc.look_at[:], desc.up[:] = info["look_from"], info["look_at"], info["up"]
    desc.vfov_degrees = info["vfov_degrees"]
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    scene.build()
    spp, depth, seed = 4, 8, 3
    ref_img, _, _ = oracle.Scene(hot, cold, mats, spheres=sp).render(oracle.camera_init(desc), spp, depth, seed=seed)
    assert ref_img.mean() > 0.01
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (im, _) in out.items():
        assert np.array_equal(im, base), mode
        assert rel_mse(im, ref_img) < 1e-4, mode
        print(f"pbrt room mode {mode}: identical to the oracle {np.mean(im == ref_img):.5f}", flush=True)
    scene.close()
