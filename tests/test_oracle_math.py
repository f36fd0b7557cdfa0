"""Pins include/mcpt_fmath.h (shared by the kernels and the CPU oracle)
against libm through numpy's float64 functions, and PCG32 against the
published pcg32 demo sequence and an independent restatement."""
import numpy as np
import pytest


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    ulp = np.maximum(ulp, np.spacing(np.float32(1e-30)))
    return np.abs(got.astype(np.float64) - ref) / ulp


@pytest.mark.parametrize("name,lo,hi,fn,bound", [
    ("sin", -30.0, 30.0, np.sin, 2.5),
    ("cos", -30.0, 30.0, np.cos, 2.5),
    ("asin", -1.0, 1.0, np.arcsin, 3.0),
    ("acos", -1.0, 1.0, np.arccos, 3.0),
    ("exp", -100.0, 88.0, np.exp, 1.5),
])
def test_unary_accuracy(oracle, name, lo, hi, fn, bound):
    x = np.linspace(lo, hi, 400001, dtype=np.float32)
    got = oracle.fmath(name, x)
    ref = fn(x.astype(np.float64))
    err = ulp_err(got, ref)
    if name in ("sin", "cos"):
        # near the zeros the reduction's absolute error (~1e-9) dominates
        err = np.where(np.abs(ref) < 1e-4, np.abs(got - ref) / 1e-9, err)
    assert err.max() <= bound, (name, float(err.max()), float(x[np.argmax(err)]))


def test_log_accuracy(oracle):
    x = np.exp(np.linspace(-100, 88, 400001)).astype(np.float32)
    x = np.concatenate([x, np.float32([1e-42, 1e-40, 1.17e-38, 0.5, 1.0, 2.0, 3.4e38])])
    got = oracle.fmath("log", x)
    ref = np.log(x.astype(np.float64))
    err = np.where(np.abs(ref) < 1e-3, np.abs(got - ref) / 1e-10, ulp_err(got, ref))
    assert err.max() <= 2.5


def test_atan2_accuracy(oracle):
    rng = np.random.default_rng(0)
    y = rng.normal(size=300000).astype(np.float32) * np.float32(3)
    x = rng.normal(size=300000).astype(np.float32) * np.float32(3)
    got = oracle.fmath("atan2", y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert ulp_err(got, ref).max() <= 4.0


def test_pow_accuracy(oracle):
    x = np.linspace(1e-6, 4.0, 200001, dtype=np.float32)
    for e in (5.0, 1 / 2.4, 2.4):
        ey = np.full_like(x, np.float32(e))
        got = oracle.fmath("pow", x, ey)
        ref = np.power(x.astype(np.float64), np.float64(np.float32(e)))
        # exp(y log x): error grows with |y log x| (documented)
        bound = 2.0 + 2.0 * np.abs(np.float64(np.float32(e)) * np.log(x.astype(np.float64)))
        assert np.all(ulp_err(got, ref) <= bound), e


def test_special_values(oracle):
    f = lambda n, *a: float(oracle.fmath(n, *[np.float32([v]) for v in a])[0])  # noqa: E731
    assert f("sin", 0.0) == 0.0 and f("cos", 0.0) == 1.0
    assert f("exp", 0.0) == 1.0 and f("log", 1.0) == 0.0
    assert f("acos", 1.0) == 0.0 and f("acos", -1.0) == np.float32(np.pi)
    assert f("asin", 1.0) == np.float32(np.pi / 2) and f("asin", -1.0) == -np.float32(np.pi / 2)
    assert f("log", 0.0) == -np.inf and np.isnan(f("log", -1.0))
    assert f("exp", -200.0) == 0.0 and f("exp", 100.0) == np.inf
    assert f("pow", 0.0, 5.0) == 0.0 and f("pow", 3.0, 0.0) == 1.0 and f("pow", 1.0, 7.0) == 1.0
    # atan2 quadrants
    assert abs(f("atan2", 1.0, 1.0) - np.pi / 4) < 1e-7
    assert abs(f("atan2", 1.0, -1.0) - 3 * np.pi / 4) < 1e-6
    assert abs(f("atan2", -1.0, -1.0) + 3 * np.pi / 4) < 1e-6
    assert abs(f("atan2", -1.0, 1.0) + np.pi / 4) < 1e-7
    assert f("atan2", 0.0, 1.0) == 0.0 and abs(f("atan2", 0.0, -1.0) - np.pi) < 1e-6
    assert abs(f("atan2", 1.0, 0.0) - np.pi / 2) < 1e-7
    # sin/cos outside the exact reduction range (|x| >= 6400) and non-finite
    # arguments give NaN, never a plausible number
    for v in (6400.0, -7000.0, 1e30, np.inf, np.nan):
        assert np.isnan(f("sin", v)) and np.isnan(f("cos", v)), v
    assert np.isfinite(f("sin", 6399.0)) and abs(f("sin", 6399.0) - np.sin(np.float64(np.float32(6399.0)))) < 1e-5


# ------------------------------------------------------------------ PCG32
MASK64 = (1 << 64) - 1


def pcg32_python(initstate, initseq, n):
    """O'Neill 2014, pcg32_srandom_r + pcg32_random_r, restated here."""
    state, inc = 0, ((initseq << 1) | 1) & MASK64
    out = []

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + initstate) & MASK64
    step()
    for _ in range(n):
        out.append(step())
    return out


def test_pcg32_published_demo():
    # pcg32-demo, pcg32_srandom_r(&rng, 42u, 54u): the published first outputs
    assert pcg32_python(42, 54, 6) == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.mark.parametrize("pixel,sample,seed", [(0, 0, 0), (1, 0, 0), (12345, 7, 1), (1048575, 63, 99),
                                               (3, 4000000, (1 << 40) + 5)])
def test_pcg32_header_matches_restatement(oracle, pixel, sample, seed):
    got = oracle.pcg32(pixel, sample, seed, 64).tolist()
    assert got == pcg32_python(((sample << 32) ^ seed) & MASK64, pixel, 64)


def test_pcg32_golden(oracle):
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "pcg32.json")
    with open(path) as f:
        g = json.load(f)
    for case in g["cases"]:
        assert oracle.pcg32(case["pixel"], case["sample"], case["seed"], len(case["out"])).tolist() == case["out"]
