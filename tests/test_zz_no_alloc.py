"""mcpt_render_noalloc: the render entry point that allocates nothing.

Called through the C ABI, its films equal mcpt_render's bit for bit on two
calls in a row (the second smaller, so it reuses the fixed workspace that
the first left behind), in every render mode, with the context's allocation
count unchanged across each call; sizes beyond the workspace are refused.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _opts(mcpt, mode, spp, depth, seed, begin=0):
    o = mcpt.RenderOpts()
    o.mode, o.spp, o.max_depth, o.seed, o.sample_begin = mode, spp, depth, seed, begin
    return o


def _used(mcpt, ctx):
    used, budget = C.c_uint64(), C.c_uint64()
    assert mcpt._lib.mcpt_ctx_memory(ctx.h, C.byref(used), C.byref(budget)) == mcpt.OK
    return used.value


CALLS = [  # (width, height, spp, depth, seed): the second smaller than the first
    (384, 288, 5, 8, 3),
    (96, 64, 3, 6, 11),
]


@pytest.mark.parametrize("mode", ["MODE_PERSISTENT", "MODE_WAVEFRONT", "MODE_MEGAKERNEL"])
def test_noalloc_matches_render(mcpt, mode):
    lib = mcpt._lib
    m = getattr(mcpt, mode)
    with mcpt.Context() as ctx:
        hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=5, param=30000)
        scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(mcpt.gen_sky(128, 64)))
        scene.build()
        films = []
        for (w, h, spp, depth, seed) in CALLS:
            print(f"noalloc: {mode} {w}x{h} spp {spp} depth {depth}", flush=True)
            cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, w, h))
            fa = ctx.film(w, h)
            before = _used(mcpt, ctx)
            o = _opts(mcpt, m, spp, depth, seed)
            assert lib.mcpt_render_noalloc(scene.h, C.byref(cam), C.byref(o), fa.h) == mcpt.OK, \
                lib.mcpt_last_error()
            assert _used(mcpt, ctx) == before  # nothing allocated from the context
            got = fa.read()
            fb = ctx.film(w, h)
            scene.render(cam, fb, mode=m, spp=spp, max_depth=depth, seed=seed)
            want = fb.read()
            assert np.array_equal(got, want), f"{mode} {w}x{h}: {np.mean(got != want)} of values differ"
            assert got.mean() > 0
            films += [fa, fb]
        for f in films:
            f.close()
        scene.close()


def test_noalloc_refuses_what_does_not_fit(mcpt):
    lib = mcpt._lib
    with mcpt.Context() as ctx:
        hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
        scene = ctx.scene(hot, cold, mats)
        scene.build()
        w, h = 640, 640  # 409,600 pixels: more paths than the workspace holds in one pass
        assert w * h > mcpt.NOALLOC_MAX_PATHS
        cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, w, h))
        film = ctx.film(w, h)
        o = _opts(mcpt, mcpt.MODE_PERSISTENT, 1, 4, 1)
        assert lib.mcpt_render_noalloc(scene.h, C.byref(cam), C.byref(o), film.h) == mcpt.ERR_INVALID
        assert b"workspace" in lib.mcpt_last_error()
        o.flags = mcpt.RENDER_TIME_KERNELS
        cam2 = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 32, 32))
        f2 = ctx.film(32, 32)
        o.mode = mcpt.MODE_WAVEFRONT
        assert lib.mcpt_render_noalloc(scene.h, C.byref(cam2), C.byref(o), f2.h) == mcpt.ERR_INVALID
        # the megakernel needs no queues: any size runs
        o = _opts(mcpt, mcpt.MODE_MEGAKERNEL, 1, 4, 1)
        assert lib.mcpt_render_noalloc(scene.h, C.byref(cam), C.byref(o), film.h) == mcpt.OK
        ref = ctx.film(w, h)
        scene.render(cam, ref, mode=mcpt.MODE_MEGAKERNEL, spp=1, max_depth=4, seed=1)
        assert np.array_equal(film.read(), ref.read())
        for f in (film, f2, ref):
            f.close()
        scene.close()
