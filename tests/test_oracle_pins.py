"""Independent pins for three parts of the CPU oracle that the renders use:
the conductor scatter (src/path_tracer.cu:520-522, :3657-3661), camera
set-up and ray generation (Camera::initialize / getRay, :1610-1658), and the
shading-normal and UV interpolation of a triangle hit (:687-713).

Every expected value is computed here from geometry in float64 (a mirror
reflection, a pinhole projection, barycentric coordinates), or from a closed
form of the sampling distribution, never by retyping the oracle's formula."""
import numpy as np
import pytest
from scipy import stats


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ---------------------------------------------------------------- conductor
def _frames(rng, k):
    d = unit(rng.normal(size=(k, 3)))
    n = unit(rng.normal(size=(k, 3)))
    n[np.sum(d * n, 1) > 0] *= -1  # the normal faces the incoming ray
    return d.astype(np.float32), n.astype(np.float32)


def test_conductor_fuzz0_is_the_mirror(oracle):
    rng = np.random.default_rng(1)
    d, n = _frames(rng, 200)
    for i in range(len(d)):
        out = oracle.conductor(d[i] * 3.7, n[i], 0.0, 4, seed=i)  # an unnormalised d is normalised first
        dd, nn = d[i].astype(np.float64), n[i].astype(np.float64)
        mirror = unit(dd - 2 * np.dot(dd, nn) * nn)
        np.testing.assert_allclose(out, np.broadcast_to(mirror, out.shape), atol=2e-6)


@pytest.mark.parametrize("fuzz", [0.1, 0.35, 0.8])
def test_conductor_fuzz_cone(oracle, fuzz):
    """|refl + fuzz*u| with |u| = 1 leaves the mirror by at most asin(fuzz)."""
    rng = np.random.default_rng(2)
    d, n = _frames(rng, 8)
    for i in range(len(d)):
        out = oracle.conductor(d[i], n[i], fuzz, 4000, seed=10 + i).astype(np.float64)
        dd, nn = d[i].astype(np.float64), n[i].astype(np.float64)
        mirror = unit(dd - 2 * np.dot(dd, nn) * nn)
        ang = np.arccos(np.clip(out @ mirror, -1, 1))
        assert ang.max() <= np.arcsin(fuzz) + 1e-4
        assert ang.max() >= 0.97 * np.arcsin(fuzz)  # the cone's edge is reached
        np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=2e-6)


def _fuzz_angle_cdf(fuzz, grid=400001):
    """u uniform on the unit sphere: c = u.r is uniform on [-1, 1], and the
    output leaves the mirror r by theta(c) = atan2(f sqrt(1-c^2), 1 + f c)."""
    c = np.linspace(-1, 1, grid)
    th = np.sort(np.arctan2(fuzz * np.sqrt(np.clip(1 - c * c, 0, None)), 1 + fuzz * c))
    return lambda x: np.searchsorted(th, x, side="right") / len(th), th


@pytest.mark.parametrize("fuzz", [0.1, 0.5])
def test_conductor_angle_and_azimuth_distribution(oracle, fuzz):
    d = np.float32([0.3, -0.8, 0.2])
    n = np.float32([0.0, 1.0, 0.0])
    N = 40000
    out = oracle.conductor(d, n, fuzz, N, seed=77).astype(np.float64)
    dd = unit(d.astype(np.float64))
    mirror = unit(dd - 2 * np.dot(dd, n) * n)
    ang = np.arccos(np.clip(out @ mirror, -1, 1))
    cdf, th = _fuzz_angle_cdf(fuzz)
    ks = stats.kstest(ang, cdf)
    assert ks.pvalue > 1e-3, ks
    # the mean cosine from the same law
    assert abs(np.cos(ang).mean() - np.cos(th).mean()) < 4 * np.cos(th).std() / np.sqrt(N) + 1e-6
    # azimuth about the mirror is uniform
    a = unit(np.cross(mirror, [1.0, 0.0, 0.0]))
    b = np.cross(mirror, a)
    phi = np.arctan2(out @ b, out @ a)
    counts, _ = np.histogram(phi, bins=16, range=(-np.pi, np.pi))
    assert stats.chisquare(counts).pvalue > 1e-3


def _quad(a, b, c, d):
    return [(a, b, c), (a, c, d)]


@pytest.mark.parametrize("fuzz", [0.0, 0.1])
def test_mirror_furnace(oracle, mcpt, fuzz):
    """A conductor plane of albedo a under a constant sky c, seen from above
    at 30-60 degrees: every path reflects once into the sky, so each sample
    is a * c (no NEE at a conductor, no MIS weight after a specular bounce)."""
    a, c = np.float32([0.9, 0.6, 0.3]), 0.7
    plane = _quad((-50, 0, 50), (50, 0, 50), (50, 0, -50), (-50, 0, -50))
    hot = np.zeros(2, oracle.HOT_DTYPE)
    cold = np.zeros(2, oracle.COLD_DTYPE)
    for i, t in enumerate(plane):
        hot[i]["v0"], hot[i]["v1"], hot[i]["v2"] = t
    mats = np.zeros(1, oracle.MAT_DTYPE)
    mats[0]["type"] = mcpt.MAT_CONDUCTOR
    mats[0]["albedo"] = a
    mats[0]["fuzz"] = fuzz
    sky = np.full((16, 32, 4), c, np.float32)
    s = oracle.Scene(hot, cold, mats, sky=sky)
    d = mcpt.CameraDesc()
    d.look_from[:] = (0, 3, 3)
    d.look_at[:] = (0, 0, 0)
    d.up[:] = (0, 1, 0)
    d.vfov_degrees, d.focus_distance, d.aperture = 25.0, 4.0, 0.0
    d.width, d.height, d.rr_cap = 32, 24, 0.95
    cam = oracle.camera_init(d)
    spp = 8
    img, _, _ = s.render(cam, spp, 8, seed=5)
    np.testing.assert_allclose(img / spp, np.broadcast_to(a * np.float32(c), img.shape), rtol=2e-6)


# ------------------------------------------------------------------ camera
def pinhole(desc, px, py):
    """The pixel-centre rays of a pinhole camera, in float64 from the
    description: the image plane at the focus distance spans
    2 tan(vfov/2) f vertically and W/H times that horizontally, row 0 at the
    top (+v), column 0 at the left (-u), and u, v, w right-handed with w
    pointing back from the view direction."""
    frm = np.array(desc.look_from[:], np.float64)
    at = np.array(desc.look_at[:], np.float64)
    up = np.array(desc.up[:], np.float64)
    W, H, f = desc.width, desc.height, desc.focus_distance
    w = unit(frm - at)
    u = unit(np.cross(up, w))
    v = np.cross(w, u)
    half_h = np.tan(np.radians(desc.vfov_degrees) / 2) * f
    half_w = half_h * W / H
    sx = (np.asarray(px) + 0.5) / W * 2 - 1  # -1 at the left edge, +1 at the right
    sy = 1 - (np.asarray(py) + 0.5) / H * 2  # +1 at the top edge
    p = frm - f * w + sx[:, None] * half_w * u + sy[:, None] * half_h * v
    return p, unit(p - frm), (u, v, w)


CAMS = [  # from, at, up, vfov, focus, W, H
    ((0, 2, 5.5), (0, 2, -2), (0, 1, 0), 40.0, 7.5, 256, 256),
    ((1.5, 0.7, -2.0), (0.2, 0.1, 0.4), (0, 1, 0), 65.0, 2.5, 1920, 1080),
    ((-3, 4, 2), (1, 0.5, -1), (0.3, 1.0, 0.45), 30.0, 1.0, 96, 64),  # rotated up vector, 3:2
    ((0, 0, 0), (0, 0, -1), (1, 0, 0), 90.0, 1.0, 33, 17),  # odd sizes: a true centre pixel
]


def _desc(mcpt, frm, at, up, vfov, focus, W, H, aperture=0.0):
    d = mcpt.CameraDesc()
    d.look_from[:] = frm
    d.look_at[:] = at
    d.up[:] = up
    d.vfov_degrees, d.focus_distance, d.aperture = vfov, focus, aperture
    d.width, d.height, d.rr_cap = W, H, 0.95
    return d


@pytest.mark.parametrize("spec", CAMS)
def test_camera_pinhole_projection(oracle, mcpt, spec):
    desc = _desc(mcpt, *spec)
    W, H = desc.width, desc.height
    cam = oracle.camera_init(desc)
    rng = np.random.default_rng(W)
    pix = np.array([(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1), (W // 2, H // 2)] +
                   [(rng.integers(W), rng.integers(H)) for _ in range(200)])
    o, dirs, times = oracle.camera_rays(cam, pix, sample=3, seed=9)
    _, expect, (u, v, w) = pinhole(desc, pix[:, 0], pix[:, 1])
    np.testing.assert_allclose(o, np.broadcast_to(np.float32(desc.look_from[:]), o.shape), atol=0)
    np.testing.assert_allclose(dirs, expect, atol=3e-6)
    assert np.all((times >= 0) & (times < 1))
    # the corner pixel's angle from the view axis: half the field of view,
    # less half a pixel, on each axis
    th = np.radians(desc.vfov_degrees) / 2
    ty = np.tan(th) * (1 - 1 / H)
    tx = np.tan(th) * W / H * (1 - 1 / W)
    np.testing.assert_allclose(np.arccos(-dirs[0].astype(np.float64) @ w), np.arctan(np.hypot(tx, ty)), atol=3e-6)
    assert dirs[0] @ v > 0 and dirs[0] @ u < 0  # top left
    if W % 2 and H % 2:
        np.testing.assert_allclose(dirs[4], -w, atol=1e-7)  # the centre pixel looks down the axis


@pytest.mark.parametrize("spec", CAMS[:3])
def test_camera_thin_lens(oracle, mcpt, spec):
    aperture = 0.6
    desc = _desc(mcpt, *spec, aperture=aperture)
    W, H = desc.width, desc.height
    cam = oracle.camera_init(desc)
    assert cam.lens_radius == np.float32(aperture / 2)
    N = 6000
    rng = np.random.default_rng(5)
    pix = np.stack([rng.integers(W, size=N), rng.integers(H, size=N)], 1)
    o, dirs, _ = oracle.camera_rays(cam, pix, sample=1, seed=4)
    p, _, (u, v, w) = pinhole(desc, pix[:, 0], pix[:, 1])
    off = o.astype(np.float64) - np.array(desc.look_from[:])
    r = np.linalg.norm(off, axis=1)
    R = aperture / 2
    assert r.max() <= R * (1 + 1e-5) and r.max() > 0.99 * R
    np.testing.assert_allclose(off @ w, 0, atol=2e-6)  # the lens lies in the plane of u and v
    # every ray through the lens still passes through its pixel's point on the focal plane
    to_p = unit(p - o.astype(np.float64))
    np.testing.assert_allclose(dirs, to_p, atol=1e-5)
    # points uniform on the disk: P(r <= x) = (x / R)^2, angle uniform
    assert stats.kstest((r / R) ** 2, "uniform").pvalue > 1e-3
    ang = np.arctan2(off @ v, off @ u)
    counts, _ = np.histogram(ang, bins=12, range=(-np.pi, np.pi))
    assert stats.chisquare(counts).pvalue > 1e-3


# ---------------------------------------------- shading normal and UV
def _tri(oracle, normals):
    hot = np.zeros(1, oracle.HOT_DTYPE)
    cold = np.zeros(1, oracle.COLD_DTYPE)
    hot[0]["v0"], hot[0]["v1"], hot[0]["v2"] = (0.2, -0.1, 0.5), (1.7, 0.4, -0.3), (-0.4, 1.3, 0.9)
    cold[0]["uv0"], cold[0]["uv1"], cold[0]["uv2"] = (0.1, 0.9), (0.8, 0.3), (0.35, 0.05)
    cold[0]["n0"], cold[0]["n1"], cold[0]["n2"] = normals
    return hot, cold


def _bary_points(hot, b):
    V = np.stack([hot[0]["v0"], hot[0]["v1"], hot[0]["v2"]]).astype(np.float64)
    return (b @ V).astype(np.float32)


def _face(hot):
    v0, v1, v2 = (hot[0][k].astype(np.float64) for k in ("v0", "v1", "v2"))
    return unit(np.cross(v1 - v0, v2 - v0))


def test_uv_is_barycentric(oracle):
    rng = np.random.default_rng(3)
    hot, cold = _tri(oracle, ((0, 0, 0), (0, 0, 0), (0, 0, 0)))
    b = rng.dirichlet((1, 1, 1), size=500)
    b = np.concatenate([np.eye(3), [[0.5, 0.5, 0], [0, 0.5, 0.5], [1 / 3, 1 / 3, 1 / 3]], b])
    uv, nrm = oracle.tri_interp(hot, cold, _bary_points(hot, b))
    UV = np.stack([cold[0]["uv0"], cold[0]["uv1"], cold[0]["uv2"]]).astype(np.float64)
    np.testing.assert_allclose(uv, b @ UV, atol=2e-5)
    # n0 = 0 means "no vertex normals": the face normal comes back unchanged
    np.testing.assert_allclose(nrm, np.broadcast_to(_face(hot), nrm.shape), atol=1e-6)


def test_shading_normal_is_interpolated(oracle):
    rng = np.random.default_rng(4)
    face = None
    ns = unit(rng.normal(size=(3, 3)))
    hot, cold = _tri(oracle, [tuple(x) for x in ns])
    face = _face(hot)
    b = np.concatenate([np.eye(3), rng.dirichlet((1, 1, 1), size=500)])
    _, nrm = oracle.tri_interp(hot, cold, _bary_points(hot, b))
    expect = unit(b @ ns)
    expect[(expect @ face) < 0] *= -1  # turned to the face's side
    np.testing.assert_allclose(nrm, expect, atol=3e-6)
    assert ((b @ ns) @ face < 0).any() and ((b @ ns) @ face > 0).any()  # both branches ran


def test_shading_normal_fallbacks(oracle):
    hot, cold = _tri(oracle, ((0, 0, 1), (0, 1, 0), (1, 0, 0)))
    face = _face(hot)
    # a point well outside the triangle: the weights sum past 1.01
    outside = _bary_points(hot, np.array([[1.4, -0.2, -0.2], [-0.5, 0.75, 0.75]]))
    _, nrm = oracle.tri_interp(hot, cold, outside)
    np.testing.assert_allclose(nrm, np.broadcast_to(face, nrm.shape), atol=1e-7)
    # opposite vertex normals cancel at the midpoint: the face normal again
    hot, cold = _tri(oracle, ((0.6, 0.0, 0.8), (-0.6, 0.0, -0.8), (0.0, 1.0, 0.0)))
    mid = _bary_points(hot, np.array([[0.5, 0.5, 0.0]]))
    _, nrm = oracle.tri_interp(hot, cold, mid)
    np.testing.assert_allclose(nrm[0], face, atol=1e-6)
