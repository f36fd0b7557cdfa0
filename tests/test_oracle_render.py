"""Pins the CPU oracle's integrator with closed forms: emitters seen
directly, a convex diffuse object under a uniform sky (L = albedo * sky),
a diffuse plane under a triangle light (Lambert's polygon formula), the
resume contract of the (pixel, sample) RNG keying, camera set-up and the
resolve curve."""
import numpy as np
import pytest


def make_materials(oracle, specs):
    m = np.zeros(len(specs), oracle.MAT_DTYPE)
    for i, (t, alb, em) in enumerate(specs):
        m[i]["type"] = t
        m[i]["albedo"] = alb
        m[i]["emission"] = em
    return m


def quad(a, b, c, d):
    return [(a, b, c), (a, c, d)]


def tris_to_arrays(oracle, tris, mat_ids):
    hot = np.zeros(len(tris), oracle.HOT_DTYPE)
    cold = np.zeros(len(tris), oracle.COLD_DTYPE)
    for i, (t, m) in enumerate(zip(tris, mat_ids)):
        hot[i]["v0"], hot[i]["v1"], hot[i]["v2"] = t
        cold[i]["material"] = m
    return hot, cold


def cam_desc(mcpt, frm, at, w, h, vfov=40.0, focus=1.0):
    d = mcpt.CameraDesc()
    d.look_from[:] = frm
    d.look_at[:] = at
    d.up[:] = (0, 1, 0) if abs(frm[1] - at[1]) < 0.99 * np.linalg.norm(np.subtract(frm, at)) else (0, 0, -1)
    d.vfov_degrees, d.focus_distance, d.aperture = vfov, focus, 0.0
    d.width, d.height, d.rr_cap = w, h, 0.95
    return d


def test_camera_matches_product(oracle, mcpt):
    for kind, w, h in ((mcpt.GEN_CORNELL, 256, 256), (mcpt.GEN_SOUP, 1024, 1024), (mcpt.GEN_SOUP, 1920, 1080),
                       (mcpt.GEN_CORNELL, 96, 64)):
        d = mcpt.gen_camera(kind, w, h)
        assert bytes(mcpt.camera_init(d)) == bytes(oracle.camera_init(d))


def test_emitter_seen_directly_is_clamped(oracle, mcpt):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, 64, 64)
    cam = oracle.camera_init(desc)
    s = oracle.Scene(hot, cold, mats)
    img, _, _ = s.render(cam, 2, 8, seed=1)
    light = np.where(mats["type"] == 3)[0][0]
    light_tris = np.where(cold["material"] == light)[0]
    # primary rays (aperture 0: origin = camera centre, no jitter)
    ys, xs = np.mgrid[0:64, 0:64]
    p = (np.array(cam.pixel00) + np.array(cam.delta_u) * xs[..., None].astype(np.float32)
         + np.array(cam.delta_v) * ys[..., None].astype(np.float32))
    d = p - np.array(cam.center)
    r = np.zeros(64 * 64, oracle.RAY_DTYPE)
    r["origin"] = np.array(cam.center)
    r["dir"] = (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)
    r["tmax"] = 1e30
    h, _, _ = s.trace_closest(r)
    sees_light = np.isin(h["prim_index"], light_tris).reshape(64, 64)
    assert sees_light.sum() > 10
    # radiance 15 seen at depth 0 has weight 1; the firefly clamp scales it to 10
    assert np.all(img[sees_light] / 2 == np.float32(10.0))


def test_convex_object_under_uniform_sky(oracle, mcpt):
    """A convex Lambertian box of albedo a in a constant sky c: every path
    leaves after one bounce, so E[L] = a * c exactly (NEE + BSDF with MIS)."""
    a, c = 0.6, 0.5
    v = [(-0.5, -0.5, -0.5), (0.5, -0.5, -0.5), (0.5, 0.5, -0.5), (-0.5, 0.5, -0.5),
         (-0.5, -0.5, 0.5), (0.5, -0.5, 0.5), (0.5, 0.5, 0.5), (-0.5, 0.5, 0.5)]
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5)]
    tris = sum((quad(*[v[i] for i in f]) for f in faces), [])
    hot, cold = tris_to_arrays(oracle, tris, [0] * len(tris))
    mats = make_materials(oracle, [(0, (a, a, a), (0, 0, 0))])
    sky = np.full((32, 64, 4), c, np.float32)
    s = oracle.Scene(hot, cold, mats, sky=sky)
    cam = oracle.camera_init(cam_desc(mcpt, (1.6, 1.2, 2.0), (0, 0, 0), 48, 48, vfov=45, focus=2.0))
    spp = 64
    img, _, _ = s.render(cam, spp, 6, seed=2)
    img = img / spp
    box = img[..., 0] < 0.45
    assert 300 < box.sum() < 48 * 48 - 300
    np.testing.assert_allclose(img[~box], c, rtol=1e-5)
    mean = img[box].mean()
    assert abs(mean - a * c) < 0.01 * a * c, mean


def lambert_polygon_irradiance(p, n, verts, L):
    """E(p) = L/2 * sum_i angle(r_i, r_i+1) * n . unit(r_i x r_i+1)"""
    r = [np.subtract(q, p) / np.linalg.norm(np.subtract(q, p)) for q in verts]
    e = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        cr = np.cross(a, b)
        e += np.arccos(np.clip(np.dot(a, b), -1, 1)) * np.dot(n, cr / np.linalg.norm(cr))
    return abs(0.5 * L * e)


def test_plane_under_triangle_light(oracle, mcpt):
    """Direct light only (the plane is convex, no sky): E[L] = a/pi * E(p)."""
    a, Le = 0.5, 4.0
    plane = quad((-20, 0, 20), (20, 0, 20), (20, 0, -20), (-20, 0, -20))
    light = [((-0.5, 1.0, -0.3), (0.6, 1.2, -0.5), (0.1, 1.1, 0.7))]
    hot, cold = tris_to_arrays(oracle, plane + light, [0, 0, 1])
    mats = make_materials(oracle, [(0, (a, a, a), (0, 0, 0)), (3, (0, 0, 0), (Le, Le, Le))])
    s = oracle.Scene(hot, cold, mats)
    desc = cam_desc(mcpt, (0.0, 3.0, 0.01), (0, 0, 0), 16, 16, vfov=50, focus=3.0)
    cam = oracle.camera_init(desc)
    spp = 2048
    img, _, _ = s.render(cam, spp, 4, seed=3)
    img = img / spp
    checked = 0
    for py, px in ((2, 3), (8, 8), (12, 5), (4, 13), (14, 14), (1, 1), (15, 8)):
        pc = np.array(cam.pixel00) + np.array(cam.delta_u) * px + np.array(cam.delta_v) * py
        d = pc - np.array(cam.center)
        t = -cam.center[1] / d[1]
        p = np.array(cam.center) + d * t
        r = np.zeros(1, oracle.RAY_DTYPE)
        r["origin"], r["dir"], r["tmax"] = np.array(cam.center), d / np.linalg.norm(d), 1e30
        if np.any(np.abs(p) > 20) or s.trace_closest(r)[0]["prim_index"][0] not in (0, 1):
            continue  # the pixel sees the light itself or misses the plane
        E = lambert_polygon_irradiance(p, np.array([0.0, 1.0, 0.0]), light[0], Le)
        expect = a / np.pi * E
        assert abs(img[py, px, 0] - expect) < 0.03 * expect + 2e-4, (py, px, img[py, px, 0], expect)
        checked += 1
    assert checked >= 4


def test_sample_ranges_resume_exactly(oracle, mcpt):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    cam = oracle.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 32, 32))
    s = oracle.Scene(hot, cold, mats)
    full, _, _ = s.render(cam, 4, 8, seed=9)
    part = np.zeros_like(full)
    s.render(cam, 2, 8, seed=9, accum=part)
    s.render(cam, 2, 8, seed=9, sample_begin=2, accum=part)
    assert np.array_equal(full, part)


def test_golden_image(oracle, mcpt):
    import os
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    img, _, _ = oracle.Scene(hot, cold, mats).render(oracle.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 64, 64)),
                                                     4, 8, seed=3)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cornell_64.npy"))
    assert np.array_equal(img, g)


@pytest.mark.parametrize("x,tm,expect", [(0.0, 0, 0.0), (1.0, 0, 1.0), (0.5, 0, 0.7353569830524495),
                                         (0.002, 0, 0.02584), (1.0, 1, 0.7353569830524495), (1e9, 2, 1.0)])
def test_resolve_curve(oracle, x, tm, expect):
    acc = np.full((1, 1, 3), x, np.float32)
    out = oracle.resolve(acc, 1, tm)
    assert abs(float(out[0, 0, 0]) - expect) < 2e-6
