"""CPU checks of the radiance-cache oracle (oracle/mlp.py).

The network has no reference counterpart, so the oracle is pinned from
outside: its backward pass against torch autograd and against central
finite differences, its encoding against closed forms, its bfloat16
rounding against hand-worked cases, and its Adam against torch.optim.Adam.
"""
import numpy as np
import pytest

import mlp as M


def _problem(n=37, seed=0):
    rng = np.random.default_rng(seed)
    feat = rng.uniform(0.0, 1.0, (n, M.FEATURES))
    targ = rng.uniform(0.0, 2.0, (n, 3))
    params = rng.normal(0.0, 0.2, M.PARAMS)
    return params, feat, targ


def test_param_count_matches_header():
    assert M.PARAMS == 4 * 64 * 64 + 3 * 64


def test_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    params, feat, targ = _problem()
    loss, grad = M.loss_and_grad(params, feat, targ, bf16_points=False)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(M.encode(feat), dtype=torch.float64)
    ws, k = [], 0
    for a, b in M.SHAPES:
        ws.append(p[k:k + a * b].reshape(a, b))
        k += a * b
    h = x
    for l in range(M.HIDDEN):
        h = torch.relu(h @ ws[l].T)
    y = h @ ws[M.HIDDEN].T
    t = torch.tensor(targ, dtype=torch.float64)
    tl = torch.mean((y - t) ** 2 / (y.detach() ** 2 + M.LOSS_EPS))
    tl.backward()
    assert abs(loss - tl.item()) < 1e-12 * max(1.0, abs(loss))
    np.testing.assert_allclose(grad, p.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_backward_matches_finite_differences():
    params, feat, targ = _problem(n=9, seed=3)
    _, grad = M.loss_and_grad(params, feat, targ, bf16_points=False)

    def loss_frozen_denominator(q):
        # the loss with the denominator held at the unperturbed output
        y0, _ = M.forward(params, feat, bf16_points=False)
        y, _ = M.forward(q, feat, bf16_points=False)
        return float(np.mean((y - targ) ** 2 / (y0 * y0 + M.LOSS_EPS)))

    rng = np.random.default_rng(5)
    idx = list(rng.choice(M.PARAMS, 24, replace=False)) + [M.PARAMS - 1, 0, 4096 * 4]
    h = 1e-6
    for i in idx:
        q1, q2 = params.copy(), params.copy()
        q1[i] += h
        q2[i] -= h
        fd = (loss_frozen_denominator(q1) - loss_frozen_denominator(q2)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])), (i, fd, grad[i])


def test_encoding_closed_forms():
    f = np.zeros((1, M.FEATURES))
    f[0, :3] = (0.25, 0.5, 0.125)
    f[0, 3:5] = (0.5, 0.25)
    f[0, 5:7] = (0.75, 1.0)
    f[0, 7] = 0.3
    f[0, 8:11] = (0.1, 0.2, 0.4)
    e = M.encode(f)[0]
    # position x = 1/4: sin/cos(pi/4), sin(pi/2) = 1, cos(pi/2) = 0, octave 2: sin(pi) = 0, cos(pi) = -1
    s2 = np.sqrt(0.5)
    np.testing.assert_allclose(e[0:6], [s2, s2, 1.0, 0.0, 0.0, -1.0], atol=1e-12)
    # position y = 1/2 (inputs 12..): sin(pi/2) = 1, cos = 0; octave 1: sin(pi) = 0, cos(pi) = -1
    np.testing.assert_allclose(e[12:16], [1.0, 0.0, 0.0, -1.0], atol=1e-12)
    # position z = 1/8, octave 3: sin(pi) = 0, cos(pi) = -1
    np.testing.assert_allclose(e[24 + 6:24 + 8], [0.0, -1.0], atol=1e-12)
    # direction (1/2, 1/4): component 0 octave 0 -> sin(pi/2), cos(pi/2); component 1 octave 1 -> sin(pi/2)
    np.testing.assert_allclose(e[36:38], [1.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(e[42:44], [1.0, 0.0], atol=1e-12)
    # normal (3/4, 1): component 1 octave 0 -> sin(pi) = 0, cos(pi) = -1
    np.testing.assert_allclose(e[48:50], [0.0, -1.0], atol=1e-12)
    np.testing.assert_allclose(e[52:64], [0.3, 0.1, 0.2, 0.4, 1.0, 0.25, 0.5, 0.125, 0.5, 0.25, 0.75, 1.0])


def test_bf16_rounding_cases():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1
    cases = {
        one + ulp / 2: one,                    # tie to even (1.0 has an even mantissa)
        one + 3 * ulp / 2: one + 2 * ulp,      # tie to even upwards
        one + ulp / 2 + 2 ** -20: one + ulp,   # above the tie
        -(one + ulp / 4): -one,
        3.0e38: 2.9999e38,                     # rounds within range
    }
    for x, want in cases.items():
        got = float(M.bf16(np.float32(x)))
        if abs(want) > 1e30:
            assert abs(got - want) / want < 1e-2
        else:
            assert got == want, (x, got, want)
    assert np.isnan(M.bf16(np.float32("nan")))


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(4)]
    ad = M.Adam(p0)
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=0.01, betas=(0.9, 0.99), eps=1e-8)
    for g in grads:
        ours = ad.step(g, 0.01)
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        np.testing.assert_allclose(ours, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_init_is_he_uniform():
    p = M.init_params(3)
    ws = M.unpack(p)
    for (a, b), w in zip(M.SHAPES, ws):
        bound = np.sqrt(6.0 / b)
        assert np.all(np.abs(w) <= bound)
        assert abs(np.std(w) - bound / np.sqrt(3.0)) < 0.1 * bound  # uniform(-b, b) has std b / sqrt(3)
    assert not np.array_equal(p, M.init_params(4))
