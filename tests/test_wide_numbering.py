"""The four-wide tree's node numbering, on the oracle's Karras trees (CPU).

The build numbers wide nodes (the binary internal nodes at even levels) in
(level, binary index) order, and the four-wide walks rely on one property of
that order: the inner children of every wide node (its binary node's inner
grandchildren, in slot order) have consecutive wide indices, so a stack pop
computes a child's index as base + rank instead of reading the parent's
child references. The property follows from Karras trees' nested key ranges;
these tests check it on random keys, on keys with long runs of duplicates
(where the index tie-break decides the splits) and on sorted runs of equal
keys, and check that a plain index-order numbering would not have it (so the
test can fail).
"""
import numpy as np
import pytest


def _levels(t):
    L, R, Ll, Rl = t["left"], t["right"], t["left_is_leaf"], t["right_is_leaf"]
    level = np.full(len(L), -1)
    level[0] = 0
    stack = [0]
    while stack:
        p = stack.pop()
        for c, leaf in ((L[p], Ll[p]), (R[p], Rl[p])):
            if not leaf:
                level[c] = level[p] + 1
                stack.append(c)
    return level


def _inner_children(t, p):
    """wide children of even-level node p that are inner, in slot order"""
    L, R, Ll, Rl = t["left"], t["right"], t["left_is_leaf"], t["right_is_leaf"]
    out = []
    for c, leaf in ((L[p], Ll[p]), (R[p], Rl[p])):
        if leaf:
            continue
        for g, gl in ((L[c], Ll[c]), (R[c], Rl[c])):
            if not gl:
                out.append(g)
    return out


def _broken_groups(t, order):
    """wide nodes whose inner children are not consecutive under `order`"""
    level = _levels(t)
    even = np.flatnonzero(level % 2 == 0)
    even = even[order(even, level)]
    widx = np.full(len(level), -1)
    widx[even] = np.arange(len(even))
    bad = 0
    for p in even:
        w = [widx[g] for g in _inner_children(t, p)]
        if w and w != list(range(w[0], w[0] + len(w))):
            bad += 1
    return bad, len(even)




def _keys(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        k = rng.integers(0, 1 << 30, n)
    elif kind == "duplicates":
        k = rng.integers(0, max(2, n // 16), n)  # runs of ~16 equal keys
    else:
        k = np.repeat(np.arange(max(1, n // 64)), 64)[:n]  # sorted runs of 64
        k = np.concatenate([k, np.full(n - len(k), k[-1] if len(k) else 0)])
    return np.sort(k.astype(np.uint32))


@pytest.mark.parametrize("kind", ["random", "duplicates", "runs"])
@pytest.mark.parametrize("n", [2, 3, 5, 17, 1000, 40000])
def test_inner_children_consecutive(oracle, kind, n):
    t = oracle.karras(_keys(kind, n, n + len(kind)))
    bad, total = _broken_groups(t, lambda even, level: np.lexsort((even, level[even])))
    assert bad == 0, (bad, total)


def test_index_order_would_break_it(oracle):
    # the numbering the walk must not be given: even-level nodes by index alone
    t = oracle.karras(_keys("random", 40000, 7))
    bad, total = _broken_groups(t, lambda even, level: np.argsort(even, kind="stable"))
    assert bad > total // 10, (bad, total)
