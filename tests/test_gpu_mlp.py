"""GPU tests of the radiance cache: the BF16 MFMA forward and backward
kernels against the CPU oracle (oracle/mlp.py, same bfloat16 rounding
points, float64 accumulation), Adam, and the cached render.

Tolerances: FP32 accumulation on the GPU against float64 in the oracle,
with a bfloat16 rounding after every layer, so an occasional activation
lands on the other side of a bfloat16 rounding boundary. Outputs: max
abs error <= 2e-2 * max |y| + 1e-3; gradients: relative L2 error <= 2e-2
per layer and cosine >= 0.999.
"""
import numpy as np
import pytest

import mlp as M
from conftest import rel_mse

pytestmark = pytest.mark.gpu


def _batch(n, seed, mcpt):
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.0, 1.0, (n, mcpt.MLP_FEATURES)).astype(np.float32)
    # a smooth radiance-like target
    t = np.stack([1.0 + np.sin(3 * f[:, 0]) * f[:, 8], 0.5 + f[:, 1] * f[:, 9], np.cos(2 * f[:, 2]) ** 2 + f[:, 5]],
                 axis=1).astype(np.float32)
    return f, t


def test_init_matches_oracle(mcpt, ctx):
    net = ctx.mlp(seed=3)
    assert np.array_equal(net.params(), M.init_params(3))
    net.close()


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4099, 100003])
def test_forward_matches_oracle(mcpt, ctx, n):
    print(f"mlp forward n={n}", flush=True)
    net = ctx.mlp(seed=5)
    p = net.params()
    f, _ = _batch(n, n, mcpt)
    y = net.forward(f)
    ref, _ = M.forward(p, f)
    err = np.abs(y - ref).max()
    assert err <= 2e-2 * np.abs(ref).max() + 1e-3, err
    net.close()


@pytest.mark.parametrize("n", [64, 777, 16384])
def test_gradient_matches_oracle(mcpt, ctx, n):
    print(f"mlp gradient n={n}", flush=True)
    net = ctx.mlp(seed=7)
    p = net.params()
    f, t = _batch(n, 100 + n, mcpt)
    loss, g = net.gradient(f, t)
    rl, rg = M.loss_and_grad(p, f, t)
    assert abs(loss - rl) <= 1e-3 * abs(rl), (loss, rl)
    k = 0
    for a, b in M.SHAPES:
        gs, rs = g[k:k + a * b], rg[k:k + a * b]
        k += a * b
        rel = np.linalg.norm(gs - rs) / np.linalg.norm(rs)
        cos = float(gs @ rs / (np.linalg.norm(gs) * np.linalg.norm(rs)))
        assert rel <= 2e-2 and cos >= 0.999, (a, b, rel, cos)
    net.close()


def test_gradient_is_deterministic(mcpt, ctx):
    net = ctx.mlp(seed=2)
    f, t = _batch(20000, 1, mcpt)
    a = net.gradient(f, t)[1]
    b = net.gradient(f, t)[1]
    assert np.array_equal(a, b)
    net.close()


def test_adam_steps_match_oracle(mcpt, ctx):
    net = ctx.mlp(seed=11)
    p0 = net.params()
    f, t = _batch(4096, 8, mcpt)
    ad = M.Adam(p0)
    for _ in range(3):
        _, rg = M.loss_and_grad(ad.p.astype(np.float32), f, t)
        ad.step(rg, 0.005)
    net.train(np.concatenate([f] * 3), np.concatenate([t] * 3), batch=4096, lr=0.005)
    got = net.params()
    # each weight moves by about lr per step; compare the moves
    move, ref_move = got - p0, ad.p - p0
    rel = np.linalg.norm(move - ref_move) / np.linalg.norm(ref_move)
    assert rel < 0.05, rel
    net.close()


def test_training_fits_a_smooth_target(mcpt, ctx):
    net = ctx.mlp(seed=1)
    f, t = _batch(65536, 21, mcpt)
    l0, _ = net.gradient(f, t)
    for _ in range(30):
        net.train(f, t, batch=16384, lr=0.005)
    l1, _ = net.gradient(f, t)
    print(f"loss {l0:.4f} -> {l1:.5f}", flush=True)
    assert l1 < 0.05 * l0
    net.close()


@pytest.mark.parametrize("frame_spp", [1, 4])
def test_untrained_frame_is_the_plain_render(mcpt, ctx, frame_spp):
    """A network that has not been trained is not queried, so the first
    frame of a cached render (frame_spp samples of every pixel traced
    together) is the plain render of those samples, bit for bit; the
    remaining samples of the call, after training, are not."""
    W, H, depth = 48, 40, 8
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, W, H))
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    net = ctx.mlp(seed=3)
    cached, plain = ctx.film(W, H), ctx.film(W, H)
    st = scene.render_cached(cam, cached, net, spp=frame_spp, sample_begin=7, max_depth=depth, seed=2,
                             train_paths=512, train_batch=1024, frame_spp=frame_spp)
    assert st["queries"] == 0 and st["paths"] == W * H * frame_spp and st["train_samples"] > 0
    scene.render(cam, plain, mode=mcpt.MODE_WAVEFRONT, spp=frame_spp, sample_begin=7, max_depth=depth, seed=2)
    assert np.array_equal(cached.read(), plain.read())
    # a second call now queries the trained network
    st = scene.render_cached(cam, cached, net, spp=2 * frame_spp, sample_begin=7 + frame_spp, max_depth=depth,
                             seed=2, train_paths=512, train_batch=1024, frame_spp=frame_spp)
    assert st["queries"] > 0 and st["paths"] == 2 * W * H * frame_spp
    for fl in (cached, plain):
        fl.close()
    net.close()
    scene.close()


def test_cached_render_gains_over_plain(mcpt, ctx):
    """Cornell box, 64x64: after the cache has trained for a few frames, a
    few cached samples per pixel are closer to a 1024-sample plain render
    than the same number of plain samples."""
    W = H = 64
    depth = 8
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, W, H))
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    ref = ctx.film(W, H)
    scene.render(cam, ref, spp=1024, max_depth=depth, seed=99)
    ref_img = ref.read() / 1024
    net = ctx.mlp(seed=1)
    warm = ctx.film(W, H)
    st = scene.render_cached(cam, warm, net, spp=32, max_depth=depth, seed=5, train_paths=1024, train_batch=4096)
    assert st["queries"] > 0 and st["train_samples"] > 0
    spp = 4
    cached = ctx.film(W, H)
    scene.render_cached(cam, cached, net, spp=spp, sample_begin=32, max_depth=depth, seed=5, train_paths=1024,
                        train_batch=4096)
    plain = ctx.film(W, H)
    scene.render(cam, plain, spp=spp, max_depth=depth, seed=5)
    c_img, p_img = cached.read() / spp, plain.read() / spp
    assert np.all(np.isfinite(c_img)) and c_img.min() >= 0.0
    e_c, e_p = rel_mse(c_img, ref_img, 0.01), rel_mse(p_img, ref_img, 0.01)
    print(f"relMSE at {spp} spp: cached {e_c:.4f} plain {e_p:.4f}", flush=True)
    assert e_c < e_p
    for fl in (ref, warm, cached, plain):
        fl.close()
    net.close()
    scene.close()
