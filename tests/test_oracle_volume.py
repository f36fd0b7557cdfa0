"""The oracle's heterogeneous medium (sampleVolumeDensity,
heterogeneousFreeFlightDistance and heterogeneousTransmittance,
src/path_tracer.cu:784-835) pinned by the mathematics, on the CPU:

  * the trilinear density reproduces a field linear in (i, j, k) exactly
    (each axis with its own slope, so a swapped axis or a wrong stride
    fails), fades to the zero background over the one-voxel shell outside
    the grid, and agrees with scipy's map_coordinates (order 1,
    grid-constant) on random grids and points;
  * in a grid of constant density equal to its maximum, delta tracking is
    the homogeneous sampler -log(1 - u) / sigma on the same first number,
    and ratio tracking is 0 or 1 on the same condition;
  * thinning: a grid at half its maximum gives free flights of rate
    sigma / 2 and transmittance exp(-sigma L / 2), not the majorant's;
  * a linear ramp: both trackers against exp(-tau) with tau the closed-form
    integral of the density along the ray.
"""
import numpy as np
import pytest

scipy_ndimage = pytest.importorskip("scipy.ndimage")


def _scene(oracle):
    sp = np.zeros(1, oracle.SPHERE_DTYPE)
    sp["radius"] = 1.0
    mats = np.zeros(1, oracle.MAT_DTYPE)
    hot, cold = np.zeros(0, oracle.HOT_DTYPE), np.zeros(0, oracle.COLD_DTYPE)
    return oracle.Scene(hot, cold, mats, spheres=sp)


def _linear_grid(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (1 + 2 * i + 4 * j + 8 * k).astype(np.float32)


def test_trilinear_reproduces_linear_field(oracle):
    rng = np.random.default_rng(1)
    dims = np.array([6, 5, 4])
    origin = np.array([0.5, -1.0, 2.0], np.float32)
    voxel = np.array([0.5, 0.25, 2.0], np.float32)
    s = _scene(oracle)
    assert s.set_density_grid(_linear_grid(*dims), origin, voxel) == 1 + 2 * 5 + 4 * 4 + 8 * 3
    q = rng.integers(0, dims - 1, (4000, 3)) + rng.integers(0, 8, (4000, 3)) / 8.0
    q = np.minimum(q, dims - 1)
    p = (origin + q * voxel).astype(np.float32)
    want = 1 + 2 * q[:, 0] + 4 * q[:, 1] + 8 * q[:, 2]
    assert np.array_equal(s.grid_density(p), want.astype(np.float32))


def test_background_shell(oracle):
    dims = (4, 3, 5)
    g = _linear_grid(*dims)
    s = _scene(oracle)
    s.set_density_grid(g, (0, 0, 0), 1.0)
    pts = np.array([[-0.5, 1, 2], [3.5, 1, 2], [1, -0.25, 2], [1, 2.75, 4], [1, 1, 4.5],
                    [-1.0, 1, 2], [4.0, 1, 2], [1, 1, -1.0], [-3e9, 0, 0], [np.nan, 1, 1], [np.inf, 1, 1]],
                   np.float32)
    want = [0.5 * g[2, 1, 0], 0.5 * g[2, 1, 3], 0.75 * g[2, 0, 1], 0.25 * g[4, 2, 1], 0.5 * g[4, 1, 1],
            0, 0, 0, 0, 0, 0]
    assert np.array_equal(s.grid_density(pts), np.array(want, np.float32))


def test_trilinear_against_scipy(oracle):
    rng = np.random.default_rng(2)
    for dims in ((7, 5, 9), (1, 4, 3), (16, 16, 16)):
        g = rng.uniform(0, 3, dims[::-1]).astype(np.float32)
        origin = rng.uniform(-2, 2, 3).astype(np.float32)
        voxel = rng.uniform(0.1, 1.0, 3).astype(np.float32)
        s = _scene(oracle)
        s.set_density_grid(g, origin, voxel)
        q = rng.uniform(-1.5, np.array(dims) + 0.5, (5000, 3))
        p = (origin + q * voxel).astype(np.float32)
        inv = (np.float32(1) / voxel).astype(np.float32)
        qf = ((p - origin).astype(np.float32) * inv).astype(np.float64)  # the sampler's own index point
        want = scipy_ndimage.map_coordinates(g.astype(np.float64), qf[:, ::-1].T, order=1, mode="grid-constant",
                                             cval=0.0)
        got = s.grid_density(p)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5), np.abs(got - want).max()


def _uniforms(oracle, n_rays, k, seed):
    """the k-th number of each ray's stream (pixel i, sample 0, seed)"""
    return np.array([(oracle.pcg32(i, 0, seed, k + 1)[k] >> 8) * 2.0 ** -24 for i in range(n_rays)], np.float32)


def test_constant_grid_is_homogeneous(oracle):
    s = _scene(oracle)
    s.set_density_grid(np.ones((8, 8, 8), np.float32), (0, 0, 0), 1.0)
    n, sigma, seed = 3000, np.float32(0.7), 5
    d = np.float32(1 / np.sqrt(3))
    rays = np.tile(np.array([1, 1, 1, d, d, d, 5.0], np.float32), (n, 1))
    u = _uniforms(oracle, n, 0, seed)
    t1 = (np.float32(0) - oracle.fmath("log", np.float32(1) - u) / sigma).astype(np.float32)
    ff = s.grid_track("free_flight", sigma, rays, seed=seed)
    assert np.array_equal(ff, np.where(t1 >= 5.0, np.float32(5.0), t1))
    tr = s.grid_track("transmittance", sigma, rays, seed=seed)
    assert np.array_equal(tr, np.where(t1 >= 5.0, 1.0, 0.0).astype(np.float32))
    assert 0.9 < (ff < 5.0).mean() < 1.0


def _within(samples, want, k=4.5):
    se = samples.std() / np.sqrt(len(samples))
    return abs(samples.mean() - want) < k * se + 1e-4


def test_thinning_uses_density_over_maximum(oracle):
    g = np.full((16, 16, 16), 0.5, np.float32)
    g[15, 15, 15] = 1.0  # sets the maximum, far from the rays
    s = _scene(oracle)
    assert s.set_density_grid(g, (0, 0, 0), 1.0) == 1.0
    n, sigma = 20000, 2.0
    rays = np.tile(np.array([1, 1, 1, 1, 0, 0, 12.0], np.float32), (n, 1))
    ff = s.grid_track("free_flight", sigma, rays, seed=9)
    assert _within(ff.astype(np.float64), 1.0 - np.exp(-12.0))  # E[min(X, 12)], X ~ Exp(sigma / 2)
    rays[:, 6] = 1.5
    tr = s.grid_track("transmittance", sigma, rays, seed=9).astype(np.float64)
    assert _within(tr, np.exp(-0.5 * sigma * 1.5))
    assert set(np.unique(tr)) <= {0.5 ** m for m in range(14)} | {0.0}


def test_linear_ramp_optical_depth(oracle):
    nx = 11
    g = np.tile((np.arange(nx) / (nx - 1)).astype(np.float32), (3, 3, 1))  # density x / 10 along x
    s = _scene(oracle)
    s.set_density_grid(g, (0, 0, 0), 1.0)
    n, sigma, L = 20000, 3.0, 4.0
    rays = np.tile(np.array([0, 1, 1, 1, 0, 0, L], np.float32), (n, 1))
    tau = lambda x: sigma * x * x / (2 * (nx - 1))
    tr = s.grid_track("transmittance", sigma, rays, seed=3).astype(np.float64)
    assert _within(tr, np.exp(-tau(L)))
    ff = s.grid_track("free_flight", sigma, rays, seed=4)
    escaped = (ff == np.float32(L)).astype(np.float64)
    assert _within(escaped, np.exp(-tau(L)))
    before = (ff < 2.0).astype(np.float64)
    assert _within(before, 1.0 - np.exp(-tau(2.0)))


def test_empty_grid_is_transparent(oracle):
    s = _scene(oracle)
    assert s.set_density_grid(np.zeros((2, 2, 2), np.float32), (0, 0, 0), 1.0) == 0.0
    rays = np.tile(np.array([0, 0, 0, 1, 0, 0, 3.0], np.float32), (10, 1))
    assert np.all(s.grid_track("free_flight", 1.0, rays) == 3.0)
    assert np.all(s.grid_track("transmittance", 1.0, rays) == 1.0)
