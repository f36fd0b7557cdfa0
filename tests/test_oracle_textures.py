"""Closed-form pins for the oracle's texture sampling (sampleTexture,
src/path_tracer.cu:216-234, with make_checker_texture's sRGB conversion,
:211-213): checker cells, and the image read's texel centres, half-way
blends, wrap in u and v, the v flip (row 0 is the top of the image) and the
sRGB byte decode."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle  # noqa: E402


def lin(e):
    return e / 12.92 if e <= 0.04045 else ((e + 0.055) / 1.055) ** 2.4


@pytest.fixture(scope="module")
def scene():
    import importlib
    mcpt = importlib.import_module("monocuda-pt_amd")
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    return pyoracle.Scene(hot, cold, mats)


def _mat(tex):
    m = np.zeros(1, pyoracle.MAT_DTYPE)[0]
    m["albedo"] = (0.3, 0.4, 0.5)
    m["texture"] = tex
    return m


def test_solid_albedo_without_texture(scene):
    out = scene.sample_albedo(_mat(0), [[0.3, 0.7], [5.0, -2.0]])
    assert np.allclose(out, [0.3, 0.4, 0.5])


def test_checker_cells(scene):
    k = scene.add_texture(1, color_a=(1.0, 0.5, 0.0), color_b=(0.0, 0.2, 1.0), scale=4.0)
    rng = np.random.default_rng(1)
    uv = rng.uniform(-2, 2, (500, 2)).astype(np.float32)
    out = scene.sample_albedo(_mat(k + 1), uv)
    even = ((np.floor(uv[:, 0] * 4).astype(int) + np.floor(uv[:, 1] * 4).astype(int)) & 1) == 0
    a = np.array([lin(1.0), lin(0.5), lin(0.0)])
    b = np.array([lin(0.0), lin(0.2), lin(1.0)])
    assert np.allclose(out[even], a, rtol=1e-6) and np.allclose(out[~even], b, rtol=1e-6)


def test_image_bilinear_wrap_and_flip(scene):
    # 2 x 2 image: top row (row 0) red, white; bottom row black, grey 128
    img = np.zeros((2, 2, 4), np.uint8)
    img[0, 0] = (255, 0, 0, 255)
    img[0, 1] = (255, 255, 255, 255)
    img[1, 0] = (0, 0, 0, 255)
    img[1, 1] = (128, 128, 128, 255)
    k = scene.add_texture(2, image=img)
    m = _mat(k + 1)
    g = lin(128 / 255)
    top, bot = np.array([[1, 0, 0], [1, 1, 1]], float), np.array([[0, 0, 0], [g, g, g]], float)
    # texel centres: u = (i + 1/2) / 2; v = 1 - (j + 1/2) / 2 picks row j
    for j, row in ((0, top), (1, bot)):
        v = 1 - (j + 0.5) / 2
        out = scene.sample_albedo(m, [[0.25, v], [0.75, v], [1.25, v], [-0.25, v]])
        assert np.allclose(out, [row[0], row[1], row[0], row[1]], atol=1e-6)
    # half-way in u (and across the wrap seam), half-way in v
    out = scene.sample_albedo(m, [[0.5, 0.75], [1.0, 0.75], [0.25, 0.5], [0.25, 1.0]])
    assert np.allclose(out[0], (top[0] + top[1]) / 2, atol=1e-6)
    assert np.allclose(out[1], (top[1] + top[0]) / 2, atol=1e-6)   # x = 1.5: texel 1 and wrapped texel 0
    assert np.allclose(out[2], (top[0] + bot[0]) / 2, atol=1e-6)   # between the rows
    assert np.allclose(out[3], (top[0] + bot[0]) / 2, atol=1e-6)   # v = 1: y = -1/2 wraps to the bottom row
    # the byte decode: 255 -> 1, 0 -> 0, 128 -> sRGB^-1(128/255)
    assert out.dtype == np.float32 and g == pytest.approx(0.2158605, abs=1e-6)
