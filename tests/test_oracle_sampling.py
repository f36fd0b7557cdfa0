"""Closed-form pins for the oracle's sampling functions: the dielectric
branch (refract, Schlick, the reflect/refract choice), the Henyey-Greenstein
sampler and pdf, and the environment-map importance sampling. Each check comes
from the mathematics (Snell's law in vector form, the HG inverse CDF, solid-
angle integrals), not from the oracle's own formula, so a dropped term, a
swapped operand or a wrong constant fails one of them.

Reference functions: refract (src/path_tracer.cu:146-163), schlick_approx
(:476-480), dielectric_scatter_direction (:524-540),
henyeyGreenstein_scatter_direction (:556-578), the HG pdf (:775-780),
environment tables and sampling (:370-385, :408-467).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle  # noqa: E402


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _incident(rng, n, cos_lo=0.05):
    """Random unit normals and unit directions with -d.n in (cos_lo, 1)."""
    nrm = _unit(rng.normal(size=(n, 3)))
    c = rng.uniform(cos_lo, 1.0, n)
    t = _unit(np.cross(nrm, rng.normal(size=(n, 3))))
    d = -c[:, None] * nrm + np.sqrt(1 - c * c)[:, None] * t
    return d.astype(np.float32), nrm.astype(np.float32), c


# ------------------------------------------------------------- dielectric
@pytest.mark.parametrize("eta", [1 / 1.5, 1.0 / 1.33, 1.5])
def test_refract_obeys_snell(eta):
    rng = np.random.default_rng(11)
    d, nrm, c = _incident(rng, 4000)
    out = pyoracle.sampling("refract", np.hstack([d, nrm, np.full((len(d), 1), eta, np.float32)])).astype(np.float64)
    sin_i = np.sqrt(1 - c * c)
    ok = eta * sin_i < 0.999  # away from the critical angle
    # Snell's law in vector form: t = eta d + (eta c - sqrt(1 - eta^2 (1 - c^2))) n
    k = 1 - eta * eta * (1 - c * c)
    t = eta * d.astype(np.float64) + (eta * c - np.sqrt(np.maximum(k, 0)))[:, None] * nrm
    assert np.abs(out[ok] - t[ok]).max() < 5e-5
    assert np.abs(np.linalg.norm(out[ok], axis=1) - 1).max() < 1e-5
    # total internal reflection: the mirror direction
    tir = eta * sin_i > 1.001
    if tir.any():
        r = d - 2 * np.sum(d * nrm, axis=1, keepdims=True) * nrm
        assert np.abs(out[tir] - r[tir]).max() < 1e-6


def test_refract_special_directions():
    n = np.array([0, 1, 0], np.float32)
    # normal incidence passes straight through whatever the ratio
    for eta in (0.5, 1.0, 1.5):
        out = pyoracle.sampling("refract", [[0, -1, 0, *n, eta]])[0]
        assert np.allclose(out, [0, -1, 0], atol=1e-6)
    # equal indices: no bending at any angle
    d, nrm, _ = _incident(np.random.default_rng(2), 500, cos_lo=0.1)
    out = pyoracle.sampling("refract", np.hstack([d, nrm, np.ones((500, 1), np.float32)]))
    assert np.abs(out - d).max() < 3e-5


@pytest.mark.parametrize("eta", [1 / 1.5, 1.5, 1.33])
def test_schlick_closed_values(eta):
    r0 = ((1 - eta) / (1 + eta)) ** 2
    c = np.linspace(0, 1, 101)
    got = pyoracle.sampling("schlick", np.stack([c, np.full_like(c, eta)], 1)).astype(np.float64)
    assert abs(got[-1] - r0) < 1e-7  # normal incidence
    assert got[0] == pytest.approx(1.0, abs=1e-7)  # grazing
    assert np.abs(got - (r0 + (1 - r0) * (1 - c) ** 5)).max() < 2e-6
    assert np.all(np.diff(got) <= 1e-7)  # monotone in the cosine
    if eta == 1.5:
        assert abs(got[-1] - 0.04) < 1e-7


def test_dielectric_choice_follows_schlick():
    """u below the Fresnel weight reflects, above it refracts; total internal
    reflection reflects for every u."""
    rng = np.random.default_rng(5)
    d, nrm, c = _incident(rng, 300, cos_lo=0.2)
    for eta in (1 / 1.5, 1.5):
        sin_i = np.sqrt(1 - c * c)
        R = pyoracle.sampling("schlick", np.stack([c, np.full_like(c, eta)], 1)).astype(np.float64)
        tir = eta * sin_i > 1.0
        R[tir] = 1.0
        mirror = d - 2 * np.sum(d * nrm, axis=1, keepdims=True) * nrm
        refr = pyoracle.sampling("refract", np.hstack([d, nrm, np.full((300, 1), eta, np.float32)]))
        for u, want_reflect in ((R * 0.999, True), (np.minimum(R + 1e-3, 0.9999), False)):
            args = np.hstack([d, nrm, np.full((300, 1), eta, np.float32), u[:, None].astype(np.float32)])
            out = pyoracle.sampling("dielectric", args)
            ref = np.abs(out - mirror).max(axis=1) < 1e-6
            if want_reflect:
                assert ref.all()
            else:
                sel = ~tir & (R + 1e-3 < 0.9999)
                assert not ref[sel].any()
                assert np.abs(out[sel] - refr[sel]).max() < 1e-6
                assert ref[tir].all()


# ------------------------------------------------------- Henyey-Greenstein
def _hg_survival(mu, g):
    """P(cos >= mu) for the HG phase function (closed-form CDF)."""
    F = (1 - g * g) / (2 * g) * ((1 + g * g - 2 * g * mu) ** -0.5 - 1 / (1 + g))
    return 1 - F


@pytest.mark.parametrize("g", [0.3, -0.5, 0.8, 0.01])
def test_hg_inverse_cdf(g):
    rng = np.random.default_rng(7)
    n = 3000
    w = _unit(rng.normal(size=(n, 3))).astype(np.float32)
    u1 = rng.uniform(0, 1, n).astype(np.float32)
    u2 = rng.uniform(0, 1, n).astype(np.float32)
    out = pyoracle.sampling("hg", np.hstack([w, np.full((n, 1), g, np.float32), u1[:, None], u2[:, None]]))
    out = out.astype(np.float64)
    assert np.abs(np.linalg.norm(out, axis=1) - 1).max() < 1e-5
    mu = np.sum(out * w, axis=1)
    # the sampled cosine sits at survival probability u1 (forward as u1 -> 0)
    assert np.abs(_hg_survival(mu, g) - u1).max() < 2e-4


def test_hg_azimuth_is_uniform_around_the_axis():
    """Averaging over a full turn of u2 leaves only the component along the
    incoming direction: the frame is orthonormal and phi = 2 pi u2."""
    w = _unit([0.3, -0.2, 0.9]).astype(np.float32)
    g, u1 = 0.4, 0.37
    u2 = (np.arange(256) + 0.5) / 256
    args = np.column_stack([np.tile(w, (256, 1)), np.full(256, g), np.full(256, u1), u2]).astype(np.float32)
    out = pyoracle.sampling("hg", args).astype(np.float64)
    mu = out @ w
    assert np.abs(mu - mu.mean()).max() < 1e-5
    assert np.abs(out.mean(axis=0) - mu.mean() * w).max() < 1e-5
    # consecutive samples are 2 pi / 256 apart around the axis
    perp = out - np.outer(mu, w)
    ang = np.arccos(np.clip(np.sum(perp[1:] * perp[:-1], 1) / np.sum(perp[1:] ** 2, 1), -1, 1))
    assert np.abs(ang - 2 * np.pi / 256).max() < 1e-3


def test_hg_isotropic_limit():
    """|g| < 1e-3 falls back to uniform sphere sampling: z = 1 - 2 u1 in the
    world frame, phi = 2 pi u2."""
    u = np.array([[0.1, 0.2], [0.5, 0.75], [0.9, 0.0]], np.float32)
    args = np.column_stack([np.tile([0, 1, 0], (3, 1)), np.full(3, 5e-4), u]).astype(np.float32)
    out = pyoracle.sampling("hg", args).astype(np.float64)
    z = 1 - 2 * u[:, 0].astype(np.float64)
    phi = 2 * np.pi * u[:, 1].astype(np.float64)
    r = np.sqrt(1 - z * z)
    assert np.abs(out - np.column_stack([r * np.cos(phi), r * np.sin(phi), z])).max() < 1e-6


@pytest.mark.parametrize("g", [0.3, -0.6, 0.9])
def test_hg_pdf_normalised_and_matches_sampler(g):
    mu = np.linspace(-1, 1, 200001)
    p = pyoracle.sampling("hg_pdf", np.stack([mu, np.full_like(mu, g)], 1)).astype(np.float64)
    dmu = mu[1] - mu[0]
    trap = lambda f: (f.sum() - 0.5 * (f[0] + f[-1])) * dmu  # noqa: E731
    assert trap(2 * np.pi * p) == pytest.approx(1.0, abs=1e-4)
    assert trap(2 * np.pi * p * mu) == pytest.approx(g, abs=1e-4)  # mean cosine is g
    # density of the sampler's cosine: -dS/dmu = 2 pi pdf
    m = np.linspace(-0.95, 0.95, 39)
    h = 1e-4
    dens = -(_hg_survival(m + h, g) - _hg_survival(m - h, g)) / (2 * h)
    pm = pyoracle.sampling("hg_pdf", np.stack([m, np.full_like(m, g)], 1)).astype(np.float64)
    assert np.abs(2 * np.pi * pm / dens - 1).max() < 1e-4


# ------------------------------------------------------------ environment
def _env_map(kind, h=16, w=32):
    rng = np.random.default_rng(3)
    if kind == "constant":
        rgb = np.ones((h, w, 3))
    elif kind == "random":
        rgb = rng.uniform(0.0, 2.0, (h, w, 3))
    else:  # one bright texel on a dim background
        rgb = np.full((h, w, 3), 0.01)
        rgb[5, 20] = 50.0
    rgba = np.concatenate([rgb, np.ones((h, w, 1))], axis=2).astype(np.float32)
    return rgba


def test_env_tables_integrate_sin_theta():
    """A constant map's total weight is w * sum_rows sin(theta_row), the
    midpoint rule for w h / pi * integral_0^pi sin = 2 w h / pi."""
    for h, w in ((16, 32), (128, 256)):
        row, col, tot = pyoracle.env_tables(_env_map("constant", h, w))
        assert tot == pytest.approx(2 * w * h / np.pi, rel=2.0 / h ** 2)
        assert np.all(np.diff(row) > 0) and row[0] == 0 and row[-1] == tot
        c = col.reshape(h, w + 1)
        # each row's running sum ends at that row's step of the row table
        # (up to the float32 rounding of the running row sum)
        assert np.abs(c[:, -1] - np.diff(row.astype(np.float64))).max() <= 4 * np.spacing(np.float32(tot))


@pytest.mark.parametrize("kind", ["constant", "random", "spot"])
@pytest.mark.parametrize("rot", [0.0, 0.7])
def test_env_pdf_integrates_to_one(kind, rot):
    rgba = _env_map(kind)
    h, w = rgba.shape[:2]
    # midpoint grid, 8x8 points per texel, d omega = sin(theta) dtheta dphi
    nt, npf = 8 * h, 8 * w
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(npf) + 0.5) / npf * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pdf = pyoracle.env_probe(rgba, rot, "pdf", d.astype(np.float32)).astype(np.float64)
    integral = (pdf * np.sin(T).ravel()).sum() * (np.pi / nt) * (2 * np.pi / npf)
    assert integral == pytest.approx(1.0, abs=3e-3)


@pytest.mark.parametrize("rot", [0.0, 1.3])
def test_env_uv_direction_inverse(rot):
    rgba = _env_map("random")
    rng = np.random.default_rng(9)
    uv = rng.uniform(0.01, 0.99, (500, 2)).astype(np.float32)
    d = pyoracle.env_probe(rgba, rot, "direction", uv)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    back = pyoracle.env_probe(rgba, rot, "uv", d)
    du = np.abs(back[:, 0] - uv[:, 0])
    du = np.minimum(du, 1 - du)  # u wraps
    assert du.max() < 2e-5 and np.abs(back[:, 1] - uv[:, 1]).max() < 2e-5
    # v = 0 is +y (theta from the pole), rotation turns about y
    top = pyoracle.env_probe(rgba, rot, "direction", [[0.3, 0.0]])[0]
    assert np.allclose(top, [0, 1, 0], atol=1e-6)


@pytest.mark.parametrize("kind", ["random", "spot"])
def test_env_sampling_is_unbiased(kind):
    """Stratified uniforms through the sampler: the returned pdf equals the
    evaluated pdf of the returned direction, and E[f / pdf] is the solid-
    angle integral of f (4 pi for f = 1; the luminance integral of the map)."""
    rgba = _env_map(kind)
    h, w = rgba.shape[:2]
    rot = 0.4
    n = 48
    g = (np.arange(n) + 0.5) / n
    j = (np.arange(8) + 0.5) / 8  # in-texel jitter
    U = np.stack(np.meshgrid(g, g, j, j, indexing="ij"), -1).reshape(-1, 4).astype(np.float32)
    out = pyoracle.env_probe(rgba, rot, "sample", U).astype(np.float64)
    d, pdf = out[:, :3], out[:, 3]
    ok = pdf > 0
    ev = pyoracle.env_probe(rgba, rot, "pdf", d[ok].astype(np.float32)).astype(np.float64)
    inner = np.abs(ev / pdf[ok] - 1) < 1e-3  # differs only on texel edges
    assert inner.mean() > 0.99
    # luminance of the texel each direction falls in
    lum = 0.2126 * rgba[..., 0] + 0.7152 * rgba[..., 1] + 0.0722 * rgba[..., 2]
    uv = pyoracle.env_probe(rgba, rot, "uv", d[ok].astype(np.float32))
    col = np.minimum((uv[:, 0] * w).astype(int), w - 1)
    row = np.minimum((uv[:, 1] * h).astype(int), h - 1)
    # positions spread uniformly inside their texel (the last two uniforms)
    for frac in (uv[:, 0] * w - col, uv[:, 1] * h - row):
        assert abs(frac.mean() - 0.5) < 0.02 and abs(frac.var() - 1 / 12) < 0.01
    f = lum[row, col].astype(np.float64)
    th_edges = np.arange(h + 1) / h * np.pi
    exact = (lum.sum(axis=1) * (2 * np.pi / w) * (np.cos(th_edges[:-1]) - np.cos(th_edges[1:]))).sum()
    est = (f / pdf[ok]).sum() / len(U)
    assert est == pytest.approx(exact, rel=2e-2)
    if kind == "spot":
        # the bright texel draws its share of the weight
        share = lum[5, 20] * np.sin(np.pi * 5.5 / h) / (lum * np.sin(np.pi * (np.arange(h)[:, None] + 0.5) / h)).sum()
        assert np.mean((row == 5) & (col == 20)) == pytest.approx(share, abs=0.02)
