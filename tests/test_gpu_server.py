"""GPU tests of the render server (mcpt_server_*).

Bar: every job's film is bit-identical to mcpt_render of the same scene,
camera, spp, depth and seed, whatever else shares the path pool, however
small the pool or the frame ring; admission follows the film budget in
FIFO order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(mcpt, ctx):
    out = []
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    out.append((ctx.scene(hot, cold, mats), mcpt.GEN_CORNELL))
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=4, param=20000)
    out.append((ctx.scene(hot, cold, mats, env=mcpt.make_env(mcpt.gen_sky(256, 128))), mcpt.GEN_SOUP))
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_MEDIUM, seed=5, param=5000)
    out.append((ctx.scene(hot, cold, mats), mcpt.GEN_SOUP))
    for sc, _ in out:
        sc.build()
    yield out
    for sc, _ in out:
        sc.close()


def _render_alone(mcpt, ctx, scene, cam, spp, depth, seed):
    film = ctx.film(cam.width, cam.height)
    scene.render(cam, film, mode=mcpt.MODE_PERSISTENT, spp=spp, max_depth=depth, seed=seed)
    img = film.read()
    film.close()
    return img


JOBS = [  # scene, width, height, spp, depth, seed
    (0, 64, 48, 5, 6, 3),
    (1, 40, 40, 3, 8, 9),
    (0, 32, 32, 7, 8, 11),
    (2, 48, 24, 4, 10, 2),
    (1, 17, 5, 9, 16, 1),
]


@pytest.mark.parametrize("capacity,frames,policy", [(0, 0, 0), (4096, 2, 0), (3100, 1, 1), (4096, 3, 2)])
def test_server_films_match_render(mcpt, ctx, scenes, capacity, frames, policy):
    print(f"server mix capacity={capacity} frames={frames} policy={policy}", flush=True)
    srv = ctx.server([sc for sc, _ in scenes], path_capacity=capacity, frames_per_job=frames, policy=policy)
    ids = []
    for (si, w, h, spp, depth, seed) in JOBS:
        cam = mcpt.camera_init(mcpt.gen_camera(scenes[si][1], w, h))
        jid, admitted = srv.submit(si, cam, spp, max_depth=depth, seed=seed)
        assert admitted
        ids.append((jid, si, cam, spp, depth, seed))
    for jid, si, cam, spp, depth, seed in ids:
        res, img = srv.wait(jid)
        assert res["status"] == mcpt.OK and res["paths"] == cam.width * cam.height * spp
        assert res["submit_ms"] <= res["start_ms"] <= res["finish_ms"]
        ref = _render_alone(mcpt, ctx, scenes[si][0], cam, spp, depth, seed)
        assert np.array_equal(img, ref), (jid, si)
        srv.release(jid)
    st = srv.stats()
    assert st["jobs_done"] == len(JOBS) and st["jobs_running"] == 0 and st["film_bytes_in_use"] == 0
    assert st["paths_started"] == sum(w * h * spp for (_, w, h, spp, _, _) in JOBS)
    srv.close()


def test_server_admission_by_budget(mcpt, ctx, scenes):
    W, H, spp = 32, 32, 4
    need = W * H * 12 * (1 + 2)  # film + two frame slots
    srv = ctx.server([scenes[0][0]], frames_per_job=2, film_budget_bytes=need + need // 2)
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, W, H))
    ids = [srv.submit(0, cam, spp, max_depth=6, seed=s) for s in (1, 2, 3)]
    assert ids[0][1] and not ids[1][1] and not ids[2][1]
    # a job larger than the whole budget is refused outright
    big = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 128, 128))
    with pytest.raises(mcpt.McptError):
        srv.submit(0, big, spp)
    with pytest.raises(mcpt.McptError):
        srv.submit(5, cam, spp)
    res0, img0 = srv.wait(ids[0][0])
    assert srv.poll(ids[1][0])["status"] == mcpt.DEFERRED  # waits for the release
    with pytest.raises(mcpt.McptError):
        srv.release(ids[1][0])  # not finished
    srv.release(ids[0][0])
    for (jid, _), seed in zip(ids[1:], (2, 3)):
        res, img = srv.wait(jid)
        assert res["start_ms"] >= res0["finish_ms"]
        assert np.array_equal(img, _render_alone(mcpt, ctx, scenes[0][0], cam, spp, 6, seed))
        srv.release(jid)
    assert np.array_equal(img0, _render_alone(mcpt, ctx, scenes[0][0], cam, spp, 6, 1))
    srv.close()
