"""The GPU path against the CPU oracle on each workload configuration's own
scenes.

- Cornell box (plain tracer): the full 256x256, 16 spp, depth 8.
- Incoherent mesh (1,048,592 triangles + sky): 262,144 closest-hit rays bit
  for bit (t, primitive, did-hit) with the reference visit order's step and
  leaf counts, any-hit, and a 128x128, 1 spp, depth-16 render.
- Batched build + render: eight icosphere scenes spanning 1k-500k triangles
  built and rendered together, each film against the oracle's render of
  that scene alone.
- Serving: four jobs of 128^2-1024^2 on resident scenes of 10k-2M triangles
  submitted together into one server, each film against the oracle.
- Radiance cache: the first frame of the cached tracer on the 2M-triangle +
  medium-box scene (the network is untrained, so the frame is the plain
  tracer's) against the oracle.
- Convergence: the GPU's FP32 film approaches the oracle's FP64-accumulated
  4096-spp image as 1/spp, and at the same seed differs from it only by
  FP32 accumulation.

Bar: relMSE = mean((a-b)^2 / (b^2 + 1e-3)) < 1e-4 for images (FP32 operation
reordering), bit equality for hits and counts.
"""
import numpy as np
import pytest

from conftest import rel_mse

pytestmark = pytest.mark.gpu


def _img_check(tag, img, ref):
    e = rel_mse(img, ref)
    same = float(np.mean(img == ref))
    print(f"{tag}: relMSE {e:.3e} identical {same:.5f}", flush=True)
    assert np.all(np.isfinite(img))
    assert e < 1e-4, f"{tag}: relMSE {e}"
    return e, same


def test_cornell_full_size(mcpt, ctx, oracle):
    W = H = 256
    spp, depth, seed = 16, 8, 1
    print("config cornell 256x256 16 spp depth 8", flush=True)
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    ref, _, ref_st = oracle.Scene(hot, cold, mats).render(oracle.camera_init(desc), spp, depth, seed=seed)
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    cam = mcpt.camera_init(desc)
    for mode in (mcpt.MODE_MEGAKERNEL, mcpt.MODE_PERSISTENT):
        film = ctx.film(W, H)
        st = scene.render(cam, film, mode=mode, spp=spp, max_depth=depth, seed=seed)
        _, same = _img_check(f"cornell mode {mode}", film.read(), ref)
        assert same == 1.0
        assert st["extend_rays"] == ref_st["extend_rays"] and st["shadow_rays"] == ref_st["shadow_rays"]
        film.close()
    scene.close()


@pytest.fixture(scope="module")
def mesh(mcpt, oracle):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=1, param=1048576)
    sky = mcpt.gen_sky(512, 256)
    return hot, cold, mats, sky, oracle.Scene(hot, cold, mats, sky=sky)


def test_mesh_closest_hit_full_size(mcpt, ctx, oracle, mesh):
    hot, cold, mats, sky, ref = mesh
    assert len(hot) == 1048592
    n = 262144
    rng = np.random.default_rng(21)
    rays = np.zeros(n, oracle.RAY_DTYPE)
    rays["origin"] = rng.uniform(-0.2, 1.2, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(1e30)
    # a quarter of them are the camera's primary rays (coherent, from outside the cube)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, 256, 256)
    o, dd, _ = oracle.camera_rays(oracle.camera_init(desc), np.stack(np.meshgrid(np.arange(256), np.arange(256)),
                                                                     -1).reshape(-1, 2))
    rays["origin"][: n // 4], rays["dir"][: n // 4] = o, dd
    print(f"config mesh: {len(hot)} tris, {n} closest-hit rays", flush=True)
    rh, rs, rl = ref.trace_closest(rays)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    hits, steps, leaves = scene.trace_closest(rays, counts=True)
    assert np.array_equal(hits["prim_index"], rh["prim_index"])
    assert np.array_equal(hits["t"].view(np.uint32), rh["t"].view(np.uint32))
    assert np.array_equal(hits["did_hit"], rh["did_hit"])
    assert np.array_equal(steps, rs) and np.array_equal(leaves, rl)
    print(f"config mesh: hit rate {rh['did_hit'].mean():.3f}, reference-order steps {rs.mean():.1f} "
          f"leaves {rl.mean():.1f} per ray", flush=True)
    assert rh["did_hit"].mean() > 0.3
    fast = scene.trace_closest(rays)  # the four-wide walk
    assert np.array_equal(fast, hits)
    rays["tmax"] = np.float32(0.02)
    assert np.array_equal(scene.trace_any(rays), ref.trace_any(rays))
    scene.close()


def test_mesh_render(mcpt, ctx, oracle, mesh):
    hot, cold, mats, sky, ref = mesh
    W = H = 128
    spp, depth, seed = 1, 16, 4
    print("config mesh render 128x128 1 spp depth 16", flush=True)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, W, H)
    ref_img, _, ref_st = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    cam = mcpt.camera_init(desc)
    for mode in (mcpt.MODE_PERSISTENT, mcpt.MODE_WAVEFRONT):
        film = ctx.film(W, H)
        st = scene.render(cam, film, mode=mode, spp=spp, max_depth=depth, seed=seed)
        _img_check(f"mesh mode {mode}", film.read(), ref_img)
        assert st["extend_rays"] == ref_st["extend_rays"]
        film.close()
    scene.close()


def test_mesh_group_walks_same_film(mcpt, ctx, mesh):
    """Short queues walked by 16 lanes per ray (the default for queues under
    32,768 rays, and forced for every bounce) give the film of one lane per
    ray bit for bit; the 128x128 render above has every bounce under the
    default threshold, so it pins the group walks against the oracle."""
    hot, cold, mats, sky, _ = mesh
    W = H = 512
    spp, depth, seed = 4, 16, 9
    print("config mesh 512x512 4 spp depth 16: group walks off / default / every bounce", flush=True)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, W, H))
    imgs, rays = [], []
    for below in (-1, 0, 1 << 30):
        film = ctx.film(W, H)
        st = scene.render(cam, film, mode=mcpt.MODE_PERSISTENT, spp=spp, max_depth=depth, seed=seed,
                          group_below=below)
        imgs.append(film.read())
        rays.append((st["extend_rays"], st["shadow_rays"]))
        film.close()
    scene.close()
    assert rays[0] == rays[1] == rays[2]
    for img in imgs[1:]:
        assert np.array_equal(img.view(np.uint32), imgs[0].view(np.uint32))


def test_batched_scenes_vs_oracle(mcpt, ctx, oracle):
    sizes = np.geomspace(1000, 500000, 8).astype(int)
    S, spp, depth, seed = 64, 8, 8, 3
    data = [mcpt.gen_scene(mcpt.GEN_ICOSPHERE, seed=200 + i, param=int(n)) for i, n in enumerate(sizes)]
    print(f"config batch: {len(data)} scenes, {[len(d[0]) for d in data]} tris", flush=True)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, S, S)
    cams = [mcpt.camera_init(desc)] * len(data)
    scenes = [ctx.scene(*d) for d in data]
    films = [ctx.film(S, S) for _ in data]
    mcpt.build_batch(scenes)
    tune, _ = mcpt.batch_autotune(scenes, cams, spp=spp, max_depth=depth, seed=seed)
    mcpt.render_batch(scenes, cams, films, spp=spp, max_depth=depth, seed=seed, tune=tune)
    rng = np.random.default_rng(5)
    rays = np.zeros(4096, oracle.RAY_DTYPE)
    rays["origin"] = rng.uniform(-1.5, 1.5, (4096, 3)).astype(np.float32)
    d = rng.normal(size=(4096, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(1e30)
    for i, (d_, sc, f) in enumerate(zip(data, scenes, films)):
        ref = oracle.Scene(*d_)
        img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
        _img_check(f"batch scene {i} ({len(d_[0])} tris)", f.read(), img)
        # the batched build's trees: reference-order visit counts pin their topology
        hits, steps, leaves = sc.trace_closest(rays, counts=True)
        rh, rs, rl = ref.trace_closest(rays)
        assert np.array_equal(hits, rh) and np.array_equal(steps, rs) and np.array_equal(leaves, rl), i
    for f in films:
        f.close()
    for sc in scenes:
        sc.close()


def test_server_jobs_vs_oracle(mcpt, ctx, oracle):
    # (triangles, side, spp): resident scenes of the serving workload's range
    jobs = [(10000, 128, 4), (150000, 256, 4), (600000, 512, 1), (2000000, 1024, 1)]
    depth = 16
    sky = mcpt.gen_sky(512, 256)
    data = [mcpt.gen_scene(mcpt.GEN_SOUP, seed=50 + i, param=n) for i, (n, _, _) in enumerate(jobs)]
    scenes = [ctx.scene(*d, env=mcpt.make_env(sky)) for d in data]
    for sc in scenes:
        sc.build()
    srv = ctx.server(scenes, policy=1)
    ids = []
    for i, (n, side, spp) in enumerate(jobs):
        cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, side, side))
        jid, admitted = srv.submit(i, cam, spp, max_depth=depth, seed=900 + i)
        assert admitted
        ids.append(jid)
    print(f"config serving: jobs {jobs} submitted together", flush=True)
    for i, ((n, side, spp), jid) in enumerate(zip(jobs, ids)):
        res, img = srv.wait(jid)
        assert res["status"] == mcpt.OK
        ref = oracle.Scene(*data[i], sky=sky)
        # the oracle covers the whole image up to 512^2 and the middle 128 rows
        # (the busiest: the camera looks at the mesh's centre) of larger ones
        r0, r1 = (0, side) if side <= 512 else (side // 2 - 64, side // 2 + 64)
        ref_img, _, _ = ref.render(oracle.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, side, side)), spp, depth,
                                   seed=900 + i, rows=(r0, r1))
        _img_check(f"server job {i} ({len(data[i][0])} tris, {side}^2, {spp} spp, rows {r0}-{r1})", img[r0:r1],
                   ref_img[r0:r1])
        srv.release(jid)
    srv.close()
    for sc in scenes:
        sc.close()


def test_cached_first_frame_vs_oracle(mcpt, ctx, oracle):
    W, H, depth, seed = 192, 108, 16, 6
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_MEDIUM, seed=1, param=2097152)
    sky = mcpt.gen_sky(512, 256)
    print(f"config cache: {len(hot)} tris + medium box, first cached frame {W}x{H}", flush=True)
    assert (mats["type"] == mcpt.MAT_ISOTROPIC).any()
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, W, H)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    net = ctx.mlp(seed=2)
    film = ctx.film(W, H)
    st = scene.render_cached(mcpt.camera_init(desc), film, net, spp=1, max_depth=depth, seed=seed,
                             train_paths=4096, train_batch=4096, frame_spp=1)
    assert st["queries"] == 0 and st["train_samples"] > 0  # untrained: no path ends in the cache
    ref_img, _, _ = oracle.Scene(hot, cold, mats, sky=sky).render(oracle.camera_init(desc), 1, depth, seed=seed)
    _img_check("cache first frame", film.read(), ref_img)
    film.close()
    net.close()
    scene.close()


def test_convergence_to_fp64_oracle(mcpt, ctx, oracle):
    """The oracle accumulates 4096 samples per pixel in FP64. The GPU's FP32
    film of other samples approaches it as 1/spp (its own variance over spp
    plus the oracle's over 4096), and the GPU film of the same 4096 samples
    differs from the FP64 sums only by FP32 accumulation."""
    W = H = 32
    depth, n_ref = 8, 4096
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    ref32, ref64, _ = oracle.Scene(hot, cold, mats).render(oracle.camera_init(desc), n_ref, depth, seed=11,
                                                          accum64=True)
    ref = ref64 / n_ref
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    cam = mcpt.camera_init(desc)
    film = ctx.film(W, H)
    scene.render(cam, film, spp=n_ref, max_depth=depth, seed=11)
    same = film.read()
    assert np.array_equal(same, ref32)  # the oracle's own FP32 sums, bit for bit
    drift = np.abs(same.astype(np.float64) - ref64) / (np.abs(ref64) + 1e-3 * n_ref)
    print(f"convergence: FP32 vs FP64 accumulation, max relative drift {drift.max():.2e}", flush=True)
    assert drift.max() < 1e-4
    film.clear()
    errs, done = [], 0
    for spp in (16, 64, 256, 1024):
        scene.render(cam, film, spp=spp - done, sample_begin=done, max_depth=depth, seed=12)
        done = spp
        errs.append(rel_mse(film.read() / spp, ref, 0.01))
    print(f"convergence: relMSE at 16/64/256/1024 spp {[f'{e:.2e}' for e in errs]}", flush=True)
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[0] / errs[-1] > 20  # 1/spp, less the oracle's own variance floor
    film.close()
    scene.close()
