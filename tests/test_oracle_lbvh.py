"""Pins the CPU oracle's LBVH stages (Morton codes, sort, Karras tree,
refit) with closed forms, an independent bit-interleave, the Karras 2012
8-key example and brute-force invariants."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def magic_expand(v):
    """10-bit spread by the usual magic-number sequence (independent of the
    loop form used by the reference and the oracle)."""
    v = np.asarray(v, np.uint64) & np.uint64(0x3FF)
    v = (v * np.uint64(0x00010001)) & np.uint64(0xFF0000FF)
    v = (v * np.uint64(0x00000101)) & np.uint64(0x0F00F00F)
    v = (v * np.uint64(0x00000011)) & np.uint64(0xC30C30C3)
    v = (v * np.uint64(0x00000005)) & np.uint64(0x49249249)
    return v.astype(np.uint32)


def test_expand_bits_closed_forms(oracle):
    lib = oracle.lib()
    assert lib.oracle_expand_bits(0x3FF) == 0x09249249
    assert lib.oracle_expand_bits(1) == 1
    assert lib.oracle_expand_bits(2) == 8
    assert lib.oracle_expand_bits(0) == 0
    allv = np.arange(1024, dtype=np.uint32)
    got = np.array([lib.oracle_expand_bits(int(v)) for v in allv], np.uint32)
    assert np.array_equal(got, magic_expand(allv))
    m = magic_expand(1023) | (magic_expand(1023) << 1) | (magic_expand(1023) << 2)
    assert int(m) == 0x3FFFFFFF


def test_quantize_boundaries(oracle):
    q = oracle.lib().oracle_quantize
    assert q(0.0, 0.0, 1.0) == 0
    assert q(np.nextafter(np.float32(1.0), np.float32(0)), 0.0, 1.0) == 1022  # truncation, no clamp
    assert q(0.5, 0.0, 1.0) == 511
    assert q(-1.0, -1.0, 3.0) == 0


def test_finalize_against_numpy(oracle, mcpt):
    hot, _, _ = mcpt.gen_scene(mcpt.GEN_SOUP, seed=4, param=5000)
    boxes, sb, cen, mort = oracle.finalize(hot)
    v = np.stack([hot["v0"], hot["v1"], hot["v2"]], 1)  # n x 3 x 3, float32
    eps = np.float32(1e-3)
    mn = v.min(1) - eps
    mx = v.max(1) + eps
    assert np.array_equal(boxes["min"], mn) and np.array_equal(boxes["max"], mx)
    assert np.array_equal(sb["min"], mn.min(0)) and np.array_equal(sb["max"], mx.max(0))
    c = ((v[:, 0] + v[:, 1]) + v[:, 2]) / np.float32(3.0)
    assert np.array_equal(cen, c)
    lo, hi = sb["min"], sb["max"]
    q = (((c - lo) / (hi - lo)) * np.float32(1023.0)).astype(np.uint32)
    expect = magic_expand(q[:, 0]) | (magic_expand(q[:, 1]) << 1) | (magic_expand(q[:, 2]) << 2)
    assert np.array_equal(mort, expect)


def test_sort_stable_and_reference(oracle):
    rng = np.random.default_rng(1)
    m = rng.integers(0, 1 << 12, 20000).astype(np.uint32)  # many duplicates
    k, i = oracle.sort_stable(m)
    order = np.argsort(m, kind="stable")
    assert np.array_equal(i, order) and np.array_equal(k, m[order])
    kr, ir = oracle.sort_reference(m)
    assert np.array_equal(kr, k)  # keys equal; ids may differ only inside equal-key runs
    assert np.array_equal(np.sort(ir), np.arange(len(m)))
    assert np.array_equal(m[ir], kr)


def test_karras_paper_example(oracle):
    with open(os.path.join(GOLDEN, "karras8.json")) as f:
        g = json.load(f)
    t = oracle.karras(np.array(g["keys"], np.uint32))
    for i, (l, ll, r, rl) in enumerate(g["nodes"]):
        assert (t["left"][i], t["left_is_leaf"][i], t["right"][i], t["right_is_leaf"][i]) == (l, ll, r, rl), i
    assert t["parent"].tolist() == g["parent"]


def _check_tree_invariants(t, n):
    L, R, LL, RL, P, LP = (t[k] for k in ("left", "right", "left_is_leaf", "right_is_leaf", "parent", "leaf_parent"))
    ni = n - 1
    # ranges by recursion from the root
    rng = {}
    stack = [(0, 0)]
    depth_max = 0
    leaves_seen = np.zeros(n, int)
    int_seen = np.zeros(ni, int)
    while stack:
        node, d = stack.pop()
        int_seen[node] += 1
        depth_max = max(depth_max, d + 1)
        for child, leaf in ((L[node], LL[node]), (R[node], RL[node])):
            if leaf:
                leaves_seen[child] += 1
                assert LP[child] == node
            else:
                assert P[child] == node
                stack.append((child, d + 1))
    assert np.all(int_seen == 1) and np.all(leaves_seen == 1)
    assert P[0] == -1
    assert depth_max <= 64
    # children of node i partition a contiguous range: leaf order is in-order
    order = []

    def walk(node):
        for child, leaf in ((L[node], LL[node]), (R[node], RL[node])):
            if leaf:
                order.append(child)
            else:
                walk(child)

    import sys
    sys.setrecursionlimit(10000)
    walk(0)
    assert order == list(range(n))


@pytest.mark.parametrize("n,bits,seed", [(2, 30, 0), (3, 30, 1), (1000, 30, 2), (5000, 6, 3), (3000, 0, 4)])
def test_karras_invariants(oracle, n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = np.sort(rng.integers(0, 1 << bits, n).astype(np.uint32)) if bits else np.zeros(n, np.uint32)
    _check_tree_invariants(oracle.karras(keys), n)


def test_refit_brute_force(oracle, mcpt):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=6, param=3000)
    s = oracle.Scene(hot, cold, mats)
    d = s.dump()
    n = len(hot)
    boxes, sb, _, _ = oracle.finalize(hot)
    lb = boxes[d["sorted_ids"]]
    assert np.array_equal(d["leaf_bounds"]["min"], lb["min"]) and np.array_equal(d["leaf_bounds"]["max"], lb["max"])
    # leaf range of every internal node, then min/max over it
    L, R, LL, RL = d["left"], d["right"], d["left_is_leaf"], d["right_is_leaf"]
    first = np.zeros(n - 1, int)
    last = np.zeros(n - 1, int)

    def rng_of(node):
        lo = L[node] if LL[node] else rng_of(L[node])[0]
        hi = R[node] if RL[node] else rng_of(R[node])[1]
        first[node], last[node] = lo, hi
        return lo, hi

    import sys
    sys.setrecursionlimit(10000)
    rng_of(0)
    for i in range(n - 1):
        seg = lb[first[i]:last[i] + 1]
        assert np.array_equal(d["internal_bounds"]["min"][i], seg["min"].min(0))
        assert np.array_equal(d["internal_bounds"]["max"][i], seg["max"].max(0))
    assert np.array_equal(d["internal_bounds"]["min"][0], sb["min"])
    assert np.array_equal(d["internal_bounds"]["max"][0], sb["max"])


def test_cornell_golden(oracle, mcpt):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    d = oracle.Scene(hot, cold, mats).dump()
    g = np.load(os.path.join(GOLDEN, "cornell_lbvh.npz"))
    for k in g.files:
        assert np.array_equal(d[k], g[k]), k
