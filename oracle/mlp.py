"""CPU oracle for the radiance-cache MLP (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import
this module; the library never calls it.

The network has no counterpart in the reference (MonoCUDA-PT has no learned
component). This file restates the network declared in include/mcpt.h
(radiance cache section) in numpy float64:

  encode    11 features -> 64 inputs (order below)
  forward   x0 = enc; x_{l+1} = relu(x_l W_l^T) for l = 0..3; y = x_4 W_4^T
  loss      mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01)
            (the relative L2 loss of Mueller et al. 2021, "Real-time Neural
            Radiance Caching for Path Tracing"; sg = no gradient)
  backward  reverse-mode through the same expressions

With bf16=True the values that the GPU rounds to bfloat16 before each
matrix product (weights, layer inputs, the output gradient and the
back-propagated gradients) are rounded here at the same points, so only
the accumulation order differs from the GPU (FP32 there, float64 here).

Parity pins: the backward pass is checked against torch autograd (float64,
bf16=False) and against central finite differences in tests/test_mlp.py;
the encoding against closed forms at chosen points.
"""
from __future__ import annotations

import numpy as np

FEATURES = 11
INPUTS = 64
WIDTH = 64
HIDDEN = 4
OUTPUTS = 3
SHAPES = [(WIDTH, INPUTS)] + [(WIDTH, WIDTH)] * (HIDDEN - 1) + [(OUTPUTS, WIDTH)]
PARAMS = sum(a * b for a, b in SHAPES)
LOSS_EPS = 0.01


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(a), a, out)
    return out.astype(np.float64)


def unpack(params):
    """Flat parameter vector -> list of weight matrices [out][in]."""
    p = np.asarray(params, np.float64)
    assert p.size == PARAMS
    ws, k = [], 0
    for a, b in SHAPES:
        ws.append(p[k:k + a * b].reshape(a, b))
        k += a * b
    return ws


def pack(ws):
    return np.concatenate([w.reshape(-1) for w in ws])


def encode(features):
    """features (n, 11) -> (n, 64). Input order:
      0..35   position axis a, octave k, sin/cos: 12a + 2k + (0 sin, 1 cos), of 2^k pi x
      36..43  direction component c: 36 + 4c + 2k + (sin, cos), octaves k < 2
      44..51  normal component c: 44 + 4c + 2k + (sin, cos)
      52 roughness, 53..55 albedo, 56 constant 1, 57..59 position,
      60..61 direction, 62..63 normal
    """
    f = np.asarray(features, np.float64)
    n = f.shape[0]
    out = np.zeros((n, INPUTS))
    for a in range(3):
        for k in range(6):
            arg = (2.0 ** k) * np.pi * f[:, a]
            out[:, 12 * a + 2 * k] = np.sin(arg)
            out[:, 12 * a + 2 * k + 1] = np.cos(arg)
    for base, off in ((36, 3), (44, 5)):
        for c in range(2):
            for k in range(2):
                arg = (2.0 ** k) * np.pi * f[:, off + c]
                out[:, base + 4 * c + 2 * k] = np.sin(arg)
                out[:, base + 4 * c + 2 * k + 1] = np.cos(arg)
    out[:, 52] = f[:, 7]
    out[:, 53:56] = f[:, 8:11]
    out[:, 56] = 1.0
    out[:, 57:60] = f[:, 0:3]
    out[:, 60:62] = f[:, 3:5]
    out[:, 62:64] = f[:, 5:7]
    return out


def forward(params, features, bf16_points=True):
    """Returns (y, cache) with y (n, 3)."""
    r = bf16 if bf16_points else (lambda v: np.asarray(v, np.float64))
    ws = [r(w) for w in unpack(params)]
    xs = [r(encode(features))]
    for l in range(HIDDEN):
        xs.append(r(np.maximum(xs[-1] @ ws[l].T, 0.0)))
    y = xs[-1] @ ws[HIDDEN].T
    return y, (ws, xs)


def loss_and_grad(params, features, targets, bf16_points=True):
    """Mean relative L2 loss and its gradient (flat, PARAMS)."""
    r = bf16 if bf16_points else (lambda v: np.asarray(v, np.float64))
    t = np.asarray(targets, np.float64)
    y, (ws, xs) = forward(params, features, bf16_points)
    n = y.shape[0]
    denom = y * y + LOSS_EPS
    loss = float(np.mean((y - t) ** 2 / denom))
    dy = r(2.0 * (y - t) / denom / (n * OUTPUTS))
    grads = [None] * (HIDDEN + 1)
    grads[HIDDEN] = dy.T @ xs[HIDDEN]
    d = dy
    for l in range(HIDDEN, 0, -1):
        dx = d @ ws[l]                       # gradient w.r.t. x_l
        d = r(dx * (xs[l] > 0))              # through the ReLU that produced x_l
        grads[l - 1] = d.T @ xs[l - 1]
    return loss, pack(grads)


def init_params(seed):
    """He-uniform weights from the library's PCG32 stream (pcg32 keyed by
    pixel = layer, sample = 0, seed): w = (2u - 1) sqrt(6 / fan_in)."""
    import pyoracle
    ws = []
    for l, (a, b) in enumerate(SHAPES):
        u = pyoracle.pcg32(l, 0, seed, a * b).astype(np.float64) * 2.0 ** -32
        u = np.asarray(u, np.float32).astype(np.float64)
        ws.append(((2.0 * u - 1.0) * np.sqrt(6.0 / b)).astype(np.float32).reshape(a, b))
    return pack(ws).astype(np.float32)


class Adam:
    """The optimiser of mcpt_mlp_train, restated (float64)."""

    def __init__(self, params, beta1=0.9, beta2=0.99, eps=1e-8):
        self.p = np.asarray(params, np.float64).copy()
        self.m = np.zeros_like(self.p)
        self.v = np.zeros_like(self.p)
        self.t = 0
        self.b1, self.b2, self.eps = beta1, beta2, eps

    def step(self, grad, lr):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * grad
        self.v = self.b2 * self.v + (1 - self.b2) * grad * grad
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.p = self.p - lr * mh / (np.sqrt(vh) + self.eps)
        return self.p
