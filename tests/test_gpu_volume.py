"""Heterogeneous media on the GPU against the CPU oracle.

A cloud in the Cornell room: an isotropic sphere that encloses a density
grid, as the reference's cloud scene wraps its NanoVDB volume
(buildCloudScene, src/path_tracer.cu:1845-1862; the grid attached to every
isotropic material, :2509-2519). Free flight inside is delta tracking and
shadow rays ratio tracking (:796-835). Renders are bit-identical across the
render modes and within relMSE 1e-4 of the oracle; the batched render, the
server and the radiance cache's first frame give the same films as the plain
render.
"""
import numpy as np
import pytest

from conftest import rel_mse
from test_gpu_parity import _render_modes

pytestmark = pytest.mark.gpu


def _cloud_grid(rng, dims=(20, 12, 16)):
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = np.zeros((nz, ny, nx), np.float32)
    for _ in range(6):  # a few soft blobs
        c = rng.uniform(0.2, 0.8, 3) * (nx, ny, nz)
        r = rng.uniform(0.15, 0.35) * nx
        d2 = ((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) / r ** 2
        g += (np.maximum(0.0, 1.0 - d2) * rng.uniform(0.5, 1.5)).astype(np.float32)
    return g


def _cloud_room(mcpt, rng, sigma=6.0):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    cloud = np.zeros(1, mcpt.MAT_DTYPE)
    cloud["type"], cloud["albedo"], cloud["density"], cloud["phase_g"] = mcpt.MAT_ISOTROPIC, 0.95, sigma, 0.6
    mats = np.concatenate([mats, cloud])
    voxel = 0.1
    grid = _cloud_grid(rng)
    extent = np.array(grid.shape[::-1], np.float32) * voxel
    center = np.array([0.0, 1.6, -2.2], np.float32)
    origin = center - extent / 2
    sp = np.zeros(1, mcpt.SPHERE_DTYPE)
    sp["center"], sp["radius"], sp["material"] = center, 0.5 * float(np.linalg.norm(extent)), len(mats) - 1
    return hot, cold, mats, sp, grid, origin, voxel


def test_render_cloud(mcpt, ctx, oracle):
    rng = np.random.default_rng(8)
    hot, cold, mats, sp, grid, origin, voxel = _cloud_room(mcpt, rng)
    W = H = 64
    spp, depth, seed = 4, 10, 23
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats, spheres=sp)
    ref = oracle.Scene(hot, cold, mats, spheres=sp)
    assert scene.set_density_grid(grid, origin, voxel) == ref.set_density_grid(grid, origin, voxel) == grid.max()
    scene.build()
    print("cloud render", flush=True)
    ref_img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    assert ref_img.mean() > 0.01
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (im, _) in out.items():
        assert np.array_equal(im, base), mode
        assert rel_mse(im, ref_img) < 1e-4, mode
        print(f"cloud mode {mode}: identical to the oracle {np.mean(im == ref_img):.5f}", flush=True)
    # the grid matters: the same room with a homogeneous medium renders otherwise
    plain = ctx.scene(hot, cold, mats, spheres=sp)
    plain.build()
    film = ctx.film(W, H)
    plain.render(cam, film, spp=spp, max_depth=depth, seed=seed)
    assert not np.array_equal(film.read(), base)
    film.close()
    plain.close()
    scene.close()


def test_cloud_batch_server_cache(mcpt, ctx):
    rng = np.random.default_rng(9)
    hot, cold, mats, sp, grid, origin, voxel = _cloud_room(mcpt, rng)
    a = ctx.scene(hot, cold, mats, spheres=sp)
    a.set_density_grid(grid, origin, voxel)
    b = ctx.scene(hot, cold, mats, spheres=sp)  # the same room, homogeneous
    soup = ctx.scene(*mcpt.gen_scene(mcpt.GEN_SOUP, seed=3, param=2000))
    scenes = [a, b, soup]
    mcpt.build_batch(scenes)
    cams = [mcpt.camera_init(mcpt.gen_camera(k, w, h)) for k, w, h in
            ((mcpt.GEN_CORNELL, 48, 40), (mcpt.GEN_CORNELL, 32, 24), (mcpt.GEN_SOUP, 24, 16))]
    spp, depth, seed = 3, 8, 4
    refs = []
    for sc, cam in zip(scenes, cams):
        f = ctx.film(cam.width, cam.height)
        sc.render(cam, f, spp=spp, max_depth=depth, seed=seed)
        refs.append(f.read())
        f.close()
    films = [ctx.film(c.width, c.height) for c in cams]
    mcpt.render_batch(scenes, cams, films, spp=spp, max_depth=depth, seed=seed)
    for i, (f, ref) in enumerate(zip(films, refs)):
        assert np.array_equal(f.read(), ref), i
        f.close()
    print("batched render with a cloud: films identical", flush=True)
    srv = ctx.server(scenes)
    ids = [srv.submit(i, cam, spp, max_depth=depth, seed=seed)[0] for i, cam in enumerate(cams)]
    for i, jid in enumerate(ids):
        res, img = srv.wait(jid)
        assert res["status"] == mcpt.OK and np.array_equal(img, refs[i]), i
        srv.release(jid)
    srv.close()
    print("server with a cloud: films identical", flush=True)
    cam = cams[0]
    net = ctx.mlp(seed=2)
    cached, plain = ctx.film(cam.width, cam.height), ctx.film(cam.width, cam.height)
    st = a.render_cached(cam, cached, net, spp=2, sample_begin=5, max_depth=8, seed=7, train_paths=256,
                         train_batch=512, frame_spp=2)
    assert st["queries"] == 0 and st["train_samples"] > 0
    a.render(cam, plain, mode=mcpt.MODE_WAVEFRONT, spp=2, sample_begin=5, max_depth=8, seed=7)
    assert np.array_equal(cached.read(), plain.read())
    print("radiance cache with a cloud: first frame identical", flush=True)
    for f in (cached, plain):
        f.close()
    net.close()
    for sc in scenes:
        sc.close()


def test_density_grid_arguments(mcpt, ctx):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    scene = ctx.scene(hot, cold, mats)
    g = np.ones((2, 2, 2), np.float32)
    for bad in (dict(density=-g), dict(density=g * np.nan), dict(voxel_size=0.0), dict(voxel_size=-1.0),
                dict(origin=(np.inf, 0, 0)), dict(density=np.ones((0, 2, 2), np.float32))):
        kw = dict(density=g, origin=(0, 0, 0), voxel_size=1.0)
        kw.update(bad)
        with pytest.raises(mcpt.McptError):
            scene.set_density_grid(**kw)
    assert scene.set_density_grid(g * 2, (0, 0, 0), 1.0) == 2.0
    scene.close()
