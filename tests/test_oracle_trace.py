"""Pins the CPU oracle's ray-triangle test and BVH traversal: known
answers on one triangle, brute force over every triangle, and the
committed closest-hit fixture."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def one_triangle(oracle, v0, v1, v2):
    hot = np.zeros(1, oracle.HOT_DTYPE)
    hot["v0"], hot["v1"], hot["v2"] = v0, v1, v2
    cold = np.zeros(1, oracle.COLD_DTYPE)
    mats = np.zeros(1, oracle.MAT_DTYPE)
    return oracle.Scene(hot, cold, mats)


def ray(o, d, tmax=1e30):
    r = np.zeros(1, np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")]))
    r["origin"], r["dir"], r["tmax"] = o, d, tmax
    return r


def test_hit_triangle_known_answers(oracle):
    s = one_triangle(oracle, (0, 0, 0), (1, 0, 0), (0, 1, 0))
    # hit at the centroid from 2 units above: t = 2 exactly
    h, st, lf = s.trace_closest(ray((1 / 3, 1 / 3, 2), (0, 0, -1)))
    assert h["did_hit"][0] == 1 and h["t"][0] == np.float32(2.0) and h["prim_index"][0] == 0
    assert st[0] == 1 and lf[0] == 2  # one root step; the single leaf sits on both sides
    # from below, direction +z: hit as well (two-sided)
    h, _, _ = s.trace_closest(ray((0.25, 0.25, -1), (0, 0, 1)))
    assert h["did_hit"][0] == 1 and h["t"][0] == np.float32(1.0)
    # parallel to the plane: |d.n| < 1e-6 -> miss
    h, _, _ = s.trace_closest(ray((0.25, 0.25, 0.0005), (1, 0, 0)))
    assert h["did_hit"][0] == 0
    # origin on the plane: t < 1e-4 -> miss
    h, _, _ = s.trace_closest(ray((0.25, 0.25, 0), (0, 0, -1)))
    assert h["did_hit"][0] == 0
    # outside each edge -> miss
    for o in ((-0.1, 0.3, 1), (0.3, -0.1, 1), (0.6, 0.6, 1)):
        h, _, _ = s.trace_closest(ray(o, (0, 0, -1)))
        assert h["did_hit"][0] == 0, o
    # beyond tmax -> miss; any-hit honours the limit too
    h, _, _ = s.trace_closest(ray((1 / 3, 1 / 3, 2), (0, 0, -1), tmax=1.5))
    assert h["did_hit"][0] == 0
    assert s.trace_any(ray((1 / 3, 1 / 3, 2), (0, 0, -1), tmax=1.5))[0] == 0
    assert s.trace_any(ray((1 / 3, 1 / 3, 2), (0, 0, -1), tmax=2.5))[0] == 1


def random_rays(n, seed, lo=-0.1, hi=1.1):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")]))
    r["origin"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = 1e30
    return r


@pytest.mark.parametrize("n_tris,seed", [(5000, 1), (40000, 2)])
def test_traversal_matches_brute_force(oracle, mcpt, n_tris, seed):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=seed, param=n_tris)
    s = oracle.Scene(hot, cold, mats)
    rays = random_rays(8192, seed)
    h, _, _ = s.trace_closest(rays)
    b = oracle.brute_closest(hot, rays)
    assert np.array_equal(h["did_hit"], b["did_hit"])
    assert np.array_equal(h["t"], b["t"])
    diff = h["prim_index"] != b["prim_index"]
    # ids may differ only on exact ties of t (first found wins)
    assert diff.mean() < 1e-3
    assert h["did_hit"].sum() > 50
    # any-hit: blocked iff the closest hit lies inside the limit
    r2 = rays.copy()
    r2["tmax"] = np.float32(0.03)
    occ = s.trace_any(r2)
    assert np.array_equal(occ.astype(bool), (b["did_hit"] == 1) & (b["t"] < np.float32(0.03)))


def test_trace_golden(oracle, mcpt):
    g = np.load(os.path.join(GOLDEN, "trace_soup.npz"))
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=5, param=4096)
    h, st, lf = oracle.Scene(hot, cold, mats).trace_closest(g["rays"])
    assert np.array_equal(h, g["hits"]) and np.array_equal(st, g["steps"]) and np.array_equal(lf, g["leaves"])
