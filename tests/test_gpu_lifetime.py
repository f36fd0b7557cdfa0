"""Handle lifetimes at the C boundary.

A context refuses to be destroyed while scenes, films, servers or networks
made from it are alive, and a scene refuses while a server renders it: the
children hold the context's stream and allocator, and a server holds views
of the scene's buffers. The calls go straight through the C ABI, as a C
caller would make them, in the wrong order first.
"""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_destroy_in_the_wrong_order_is_an_error(mcpt):
    lib = mcpt._lib
    VP = C.c_void_p
    print("lifetime: raw C-ABI handles, wrong destroy order", flush=True)
    ctx = VP()
    assert lib.mcpt_ctx_create(0, 0, C.byref(ctx)) == mcpt.OK
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    scene = VP()
    assert lib.mcpt_scene_create(ctx, mcpt._ptr(hot), mcpt._ptr(cold), len(hot), mcpt._ptr(mats), len(mats), None,
                                 C.byref(scene)) == mcpt.OK
    assert lib.mcpt_lbvh_build(scene, None) == mcpt.OK
    film = VP()
    assert lib.mcpt_film_create(ctx, 8, 8, C.byref(film)) == mcpt.OK
    mlp = VP()
    assert lib.mcpt_mlp_create(ctx, 3, C.byref(mlp)) == mcpt.OK
    srv = VP()
    arr = (VP * 1)(scene)
    opts = mcpt.ServerOpts(4096, 2, 0, 0, 0, 0, 0)
    assert lib.mcpt_server_create(ctx, arr, 1, C.byref(opts), C.byref(srv)) == mcpt.OK

    # children alive: the context stays, and says why
    assert lib.mcpt_ctx_destroy(ctx) == mcpt.ERR_INVALID
    assert b"live" in lib.mcpt_last_error()
    # a server renders the scene: the scene stays
    assert lib.mcpt_scene_destroy(scene) == mcpt.ERR_INVALID
    assert b"server" in lib.mcpt_last_error()
    # everything still works after the refused calls
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 8, 8))
    opts_r = mcpt.RenderOpts()
    opts_r.mode = mcpt.MODE_WAVEFRONT
    opts_r.spp = 1
    opts_r.max_depth = 4
    assert lib.mcpt_render(scene, C.byref(cam), C.byref(opts_r), film, None) == mcpt.OK

    assert lib.mcpt_server_destroy(srv) == mcpt.OK
    assert lib.mcpt_scene_destroy(scene) == mcpt.OK
    assert lib.mcpt_ctx_destroy(ctx) == mcpt.ERR_INVALID  # film and network remain
    assert lib.mcpt_film_destroy(film) == mcpt.OK
    assert lib.mcpt_ctx_destroy(ctx) == mcpt.ERR_INVALID  # the network remains
    assert lib.mcpt_mlp_destroy(mlp) == mcpt.OK
    assert lib.mcpt_ctx_destroy(ctx) == mcpt.OK


def test_python_close_order(mcpt):
    """The binding closes a context's children first; closing a scene that a
    live server renders raises instead of freeing it."""
    ctx = mcpt.Context()
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    sc = ctx.scene(hot, cold, mats)
    sc.build()
    srv = ctx.server([sc], path_capacity=4096)
    with pytest.raises(mcpt.McptError) as e:
        sc.close()
    assert e.value.code == mcpt.ERR_INVALID
    ctx.close()  # server, then scene, then the context
    assert not srv.h and not sc.h and not ctx.h
