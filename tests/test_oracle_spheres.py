"""The oracle's spheres (sphere / hit_sphere / computeSphereUvFromOutwardNormal
and finalizeScene's primitive list, src/path_tracer.cu:845-911, :1930-1972)
pinned by the mathematics, on the CPU:

  * closest hits in mixed sphere + triangle scenes against a float64 brute
    force that solves |o + t d - c|^2 = r^2 and the triangle plane exactly;
  * a ray from inside a sphere leaves through the far root;
  * a moving sphere is where c0 + (c1 - c0) t puts it at the ray's time;
  * the sphere uv against its closed form;
  * the build's primitive list: spheres first, their boxes centre +- radius
    (the union over the motion for a moving one), the scene box the exact
    min/max over every primitive box.
"""
import numpy as np
import pytest


def _spheres(oracle, centers, radii, moving=None, c1=None, material=0):
    sp = np.zeros(len(centers), oracle.SPHERE_DTYPE)
    sp["center"] = centers
    sp["radius"] = radii
    sp["material"] = material
    if moving is not None:
        sp["moving"] = moving
        sp["center1"] = c1
    return sp


def _mixed(oracle, rng, n_sph=40, n_tri=40):
    c = rng.uniform(-4, 4, (n_sph, 3)).astype(np.float32)
    r = rng.uniform(0.1, 0.6, n_sph).astype(np.float32)
    sp = _spheres(oracle, c, r)
    base = rng.uniform(-4, 4, (n_tri, 1, 3))
    v = (base + rng.uniform(-0.7, 0.7, (n_tri, 3, 3))).astype(np.float32)
    hot = np.zeros(n_tri, oracle.HOT_DTYPE)
    hot["v0"], hot["v1"], hot["v2"] = v[:, 0], v[:, 1], v[:, 2]
    cold = np.zeros(n_tri, oracle.COLD_DTYPE)
    mats = np.zeros(1, oracle.MAT_DTYPE)
    mats["albedo"] = 0.5
    return hot, cold, mats, sp


def _rays(rng, n, time=None):
    rays = np.zeros(n, np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")]))
    o = rng.uniform(-6, 6, (n, 3))
    tgt = rng.uniform(-3, 3, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["origin"], rays["dir"], rays["tmax"] = o, d, 1e30
    rays["time"] = rng.uniform(0, 1, n) if time is None else time
    return rays


def _brute(hot, sp, rays):
    """float64 closest hit over spheres (type 0) and triangles (type 1):
    (t, type, index, gap to the second-nearest candidate)"""
    o = rays["origin"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    n = len(rays)
    cand = []
    for k in range(len(sp)):
        c0 = sp["center"][k].astype(np.float64)
        c = c0 + (sp["center1"][k] - c0) * rays["time"][:, None] if sp["moving"][k] else c0[None, :]
        oc = o - c
        b = np.einsum("ij,ij->i", oc, d)
        q = np.einsum("ij,ij->i", oc, oc) - float(sp["radius"][k]) ** 2
        disc = b * b - q
        h = np.sqrt(np.maximum(disc, 0))
        t = np.where(-b - h > 1e-4, -b - h, -b + h)
        t = np.where((disc >= 0) & (t > 1e-4), t, np.inf)
        cand.append((t, 0, k))
    for k in range(len(hot)):
        v0, v1, v2 = (hot[f][k].astype(np.float64) for f in ("v0", "v1", "v2"))
        e1, e2 = v1 - v0, v2 - v0
        nrm = np.cross(e1, e2)
        nrm /= np.linalg.norm(nrm)
        den = d @ nrm
        t = ((v0 - o) @ nrm) / np.where(np.abs(den) < 1e-12, 1e-12, den)
        p = o + d * t[:, None]
        ok = (np.abs(den) >= 1e-6) & (t > 1e-4)
        for a, b_ in ((v0, v1), (v1, v2), (v2, v0)):
            ok &= np.cross(b_ - a, p - a) @ nrm >= 0
        cand.append((np.where(ok, t, np.inf), 1, k))
    ts = np.stack([c[0] for c in cand])
    order = np.argsort(ts, axis=0)
    best = order[0]
    t_best = ts[best, np.arange(n)]
    t_second = ts[order[1], np.arange(n)] if len(cand) > 1 else np.full(n, np.inf)
    types = np.array([c[1] for c in cand])[best]
    idx = np.array([c[2] for c in cand])[best]
    with np.errstate(invalid="ignore"):
        gap = np.where(np.isinf(t_best), np.inf, t_second - t_best)
    return t_best, types, idx, gap


def test_closest_hits_mixed_scene(oracle):
    rng = np.random.default_rng(11)
    hot, cold, mats, sp = _mixed(oracle, rng)
    scene = oracle.Scene(hot, cold, mats, spheres=sp)
    rays = _rays(rng, 4000)
    hits, _, _ = scene.trace_closest(rays)
    t, ty, idx, gap = _brute(hot, sp, rays)
    hit = np.isfinite(t)
    assert hit.mean() > 0.3 and (ty[hit] == 0).sum() > 200 and (ty[hit] == 1).sum() > 200
    # grazing rays (a near-tangent sphere hit, a hit on a triangle edge) can
    # go either way in float32; everything else agrees
    clear = gap > 1e-3 * np.maximum(t, 1)
    assert np.array_equal(hits["did_hit"][clear] == 1, hit[clear])
    both = hit & clear & (hits["did_hit"] == 1)
    assert np.array_equal(hits["prim_type"][both], ty[both])
    assert np.array_equal(hits["prim_index"][both], idx[both])
    assert np.allclose(hits["t"][both], t[both], rtol=2e-5, atol=1e-5)
    assert np.all(hits["prim_type"][hits["did_hit"] == 0] == -1)
    # any-hit agrees with the closest hit
    occ = scene.trace_any(rays)
    assert np.array_equal(occ == 1, hits["did_hit"] == 1)


def test_ray_from_inside_leaves_through_far_root(oracle):
    mats = np.zeros(1, oracle.MAT_DTYPE)
    sp = _spheres(oracle, [[1.0, 2.0, 3.0]], [2.0])
    scene = oracle.Scene(np.zeros(0, oracle.HOT_DTYPE), np.zeros(0, oracle.COLD_DTYPE), mats, spheres=sp)
    rays = np.zeros(4, oracle.RAY_DTYPE)
    rays["origin"] = [1.0, 2.0, 3.0]
    rays["dir"] = [[1, 0, 0], [0, -1, 0], [0, 0, 1], [0.6, 0.8, 0]]
    rays["tmax"] = 1e30
    hits, _, _ = scene.trace_closest(rays)
    assert np.all(hits["did_hit"] == 1) and np.all(hits["prim_type"] == 0)
    assert np.allclose(hits["t"], 2.0, rtol=1e-6)
    # starting outside, facing away: no hit; tmax below the entry: no hit
    rays["origin"] = [10.0, 2.0, 3.0]
    rays["dir"] = [1, 0, 0]
    assert np.all(scene.trace_closest(rays)[0]["did_hit"] == 0)
    rays["dir"] = [-1, 0, 0]
    rays["tmax"] = 6.9
    assert np.all(scene.trace_any(rays) == 0)
    rays["tmax"] = 7.1
    assert np.all(scene.trace_any(rays) == 1)


def test_moving_sphere_follows_time(oracle):
    mats = np.zeros(1, oracle.MAT_DTYPE)
    sp = _spheres(oracle, [[0.0, 0.0, 0.0]], [0.5], moving=[1], c1=[[4.0, 0.0, 0.0]])
    scene = oracle.Scene(np.zeros(0, oracle.HOT_DTYPE), np.zeros(0, oracle.COLD_DTYPE), mats, spheres=sp)
    times = np.array([0.0, 0.25, 0.5, 0.75, 0.999], np.float32)
    rays = np.zeros(len(times), oracle.RAY_DTYPE)
    rays["origin"][:, 0] = 4.0 * times   # straight down onto where the centre is at that time
    rays["origin"][:, 1] = 10.0
    rays["dir"] = [0, -1, 0]
    rays["tmax"] = 1e30
    rays["time"] = times
    hits, _, _ = scene.trace_closest(rays)
    assert np.all(hits["did_hit"] == 1) and np.allclose(hits["t"], 9.5, rtol=1e-6)
    # the same rays at time 0 miss everywhere but over the start
    rays["time"] = 0.0
    hits, _, _ = scene.trace_closest(rays)
    assert list(hits["did_hit"]) == [1, 0, 0, 0, 0]
    # the box covers the whole motion
    d = scene.dump()
    lb = d["leaf_bounds"][0]
    assert np.allclose(lb["min"], [-0.5, -0.5, -0.5]) and np.allclose(lb["max"], [4.5, 0.5, 0.5])


def test_sphere_uv_closed_form(oracle):
    rng = np.random.default_rng(2)
    n = rng.normal(size=(2000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = n.astype(np.float32)
    uv = oracle.sphere_uv(n).astype(np.float64)
    x, y, z = (n[:, k].astype(np.float64) for k in range(3))
    u = (np.arctan2(-z, x) + np.pi) / (2 * np.pi)
    v = np.arccos(np.clip(-y, -1, 1)) / np.pi
    assert np.allclose(uv[:, 0], u, atol=2e-6) and np.allclose(uv[:, 1], v, atol=2e-6)
    # the poles and the seam: -y at v = 0, +y at v = 1, +x at u = 1/2
    pts = oracle.sphere_uv(np.array([[0, -1, 0], [0, 1, 0], [1, 0, 0], [0, 0, -1]], np.float32))
    assert pts[0, 1] == 0.0 and abs(pts[1, 1] - 1.0) < 1e-6 and abs(pts[2, 0] - 0.5) < 1e-6
    assert abs(pts[3, 0] - 0.75) < 1e-6


def test_primitive_list_and_boxes(oracle):
    rng = np.random.default_rng(5)
    hot, cold, mats, sp = _mixed(oracle, rng, n_sph=30, n_tri=50)
    sp["moving"][::3] = 1
    sp["center1"][::3] = sp["center"][::3] + rng.uniform(-1, 1, (10, 3)).astype(np.float32)
    scene = oracle.Scene(hot, cold, mats, spheres=sp)
    d = scene.dump()
    S, T = len(sp), len(hot)
    ids = d["sorted_ids"]
    assert sorted(ids) == list(range(S + T))
    r = sp["radius"].astype(np.float32)[:, None]
    smin = np.where(sp["moving"][:, None] == 1, np.minimum(sp["center"] - r, sp["center1"] - r), sp["center"] - r)
    smax = np.where(sp["moving"][:, None] == 1, np.maximum(sp["center"] + r, sp["center1"] + r), sp["center"] + r)
    v = np.stack([hot["v0"], hot["v1"], hot["v2"]], 1)
    tmin, tmax = v.min(1) - np.float32(1e-3), v.max(1) + np.float32(1e-3)
    pmin, pmax = np.concatenate([smin, tmin]), np.concatenate([smax, tmax])
    assert np.array_equal(d["leaf_bounds"]["min"], pmin[ids]) and np.array_equal(d["leaf_bounds"]["max"], pmax[ids])
    assert np.array_equal(d["scene_bounds"]["min"][0], pmin.min(0))
    assert np.array_equal(d["scene_bounds"]["max"][0], pmax.max(0))
    # codes are sorted and each parent box holds its children
    assert np.all(np.diff(d["sorted_keys"].astype(np.int64)) >= 0)
    ib = d["internal_bounds"]
    for i in range(S + T - 1):
        for child, leaf in ((d["left"][i], d["left_is_leaf"][i]), (d["right"][i], d["right_is_leaf"][i])):
            cb = d["leaf_bounds"][child] if leaf else ib[child]
            assert np.all(ib[i]["min"] <= cb["min"]) and np.all(ib[i]["max"] >= cb["max"])
