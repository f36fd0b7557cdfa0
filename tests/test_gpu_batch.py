"""GPU tests of the batched scenes (mcpt_lbvh_build_batch, mcpt_render_batch,
mcpt_batch_autotune): every scene's BVH must traverse exactly like the one
its own mcpt_lbvh_build makes (reference-order step and leaf counts and hits,
bit for bit), and every film must equal mcpt_render of its scene bit for bit,
for any tuning.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 17, 1000, 4097, 30000]


def _scene_data(mcpt, n, seed):
    if n <= 16:
        hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=seed, param=16)
        return hot[:n], cold[:n], mats
    kind = mcpt.GEN_ICOSPHERE if seed % 2 else mcpt.GEN_SOUP
    return mcpt.gen_scene(kind, seed=seed, param=n)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")]))
    r["origin"] = rng.uniform(-0.3, 1.3, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmax"] = 1e30
    return r


def test_batched_build_matches_single_builds(mcpt, ctx):
    data = [_scene_data(mcpt, n, 40 + i) for i, n in enumerate(SIZES)]
    single = [ctx.scene(*d) for d in data]
    for sc in single:
        sc.build()
    batched = [ctx.scene(*d) for d in data]
    ms = mcpt.build_batch(batched)
    print(f"batched build of {len(SIZES)} scenes: {ms:.3f} ms", flush=True)
    rays = _rays(20000, 3)
    for n, a, b in zip(SIZES, single, batched):
        ha, sa, la = a.trace_closest(rays, counts=True)
        hb, sb, lb = b.trace_closest(rays, counts=True)
        assert np.array_equal(ha, hb) and np.array_equal(sa, sb) and np.array_equal(la, lb), n
        assert np.array_equal(a.trace_closest(rays), b.trace_closest(rays)), n
        # the four-wide tree and leaf ranks are the single build's, bit for bit
        wa, ta = a.dump_wide()
        wb, tb = b.dump_wide()
        assert wa.tobytes() == wb.tobytes() and np.array_equal(ta, tb), n
    for sc in single + batched:
        sc.close()


def test_batched_render_matches_single_renders(mcpt, ctx):
    sizes = [3000, 12000, 500, 40000]
    data = [_scene_data(mcpt, n, 60 + i) for i, n in enumerate(sizes)]
    # and a scene with a homogeneous medium (its shadow rays take the
    # transmittance walk inside the persistent shadow kernel)
    data.append(mcpt.gen_scene(mcpt.GEN_MEDIUM, seed=65, param=20000))
    scenes = [ctx.scene(*d) for d in data]
    mcpt.build_batch(scenes)
    dims = [(40, 30), (64, 64), (17, 9), (48, 48), (32, 24)]
    cams = [mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, w, h)) for w, h in dims]
    spp, depth, seed = 5, 8, 11
    refs = []
    for sc, cam in zip(scenes, cams):
        f = ctx.film(cam.width, cam.height)
        sc.render(cam, f, spp=spp, max_depth=depth, seed=seed)
        refs.append(f.read())
        f.close()
    tunes = [None, mcpt.BatchTune(1, 16, 0, 0), mcpt.BatchTune(3, 48, 2, 0), mcpt.BatchTune(8, 32, 0, 0)]
    for tune in tunes:
        print(f"batched render tune={tune.as_dict() if tune else None}", flush=True)
        films = [ctx.film(c.width, c.height) for c in cams]
        st = mcpt.render_batch(scenes, cams, films, spp=spp, max_depth=depth, seed=seed, tune=tune)
        assert st["paths"] == sum(w * h for w, h in dims) * spp
        for i, (f, ref) in enumerate(zip(films, refs)):
            assert np.array_equal(f.read(), ref), (i, tune)
            f.close()
    best, search_ms = mcpt.batch_autotune(scenes, cams, spp=spp, max_depth=depth, seed=seed)
    print(f"autotune {best.as_dict()} in {search_ms:.1f} ms", flush=True)
    assert best.samples_per_pass >= 1 and best.refill_lanes in (16, 32, 48) and best.blocks_per_cu in (0, 2, 4)
    # depth 1: the timed launches trace the camera rays themselves
    one, one_ms = mcpt.batch_autotune(scenes, cams, spp=1, max_depth=1, seed=seed)
    assert one.samples_per_pass == 1 and one.refill_lanes in (16, 32, 48) and one_ms > 0.0
    films = [ctx.film(c.width, c.height) for c in cams]
    mcpt.render_batch(scenes, cams, films, spp=spp, max_depth=depth, seed=seed, tune=best)
    for f, ref in zip(films, refs):
        assert np.array_equal(f.read(), ref)
        f.close()
    for sc in scenes:
        sc.close()
