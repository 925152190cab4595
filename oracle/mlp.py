"""FP64 reference of the surrogate MLP (forward, loss, backward, Adam).

TEST INFRASTRUCTURE ONLY: the checker for the GPU training and inference
kernels (tests/, bench.py's error judge); the product never imports it.

The network and its mixed-precision data path are the project's own (the
reference has no learned component): 64 -> 64 -> 64 -> 64 -> 64 with ReLU on
the hidden layers; inputs per direction (log(1 + c) rgb, cone transmittance)
x 16; outputs 0..47 are the level-2 far-field residual (16 directions x rgb,
normalised), output 48 the log of the predicted squared
error, outputs 49..63 padding.  Layer l computes z = bf16(a) . bf16(W) + b
with FP32 accumulation on the GPU; this module evaluates the same rounding
points in FP64, so the GPU differs from it only by accumulation order.
`emulate_bf16=False` gives the exact FP64 network (pinned by finite
differences and by torch autograd in tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT_PAD = 64, 64, 64
N_RAD = 48          # radiance outputs
ERR = 48            # index of the error head
SIZES = [(IN, HID), (HID, HID), (HID, HID), (HID, OUT_PAD)]
ERR_WEIGHT = 0.05   # weight of the error-head loss
ERR_EPS = 1e-2


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even, via fp32) and back to float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int = 0):
    """He-uniform weights [Kin, Nout], zero biases; padding outputs zero."""
    rng = np.random.default_rng(seed)
    params = []
    for l, (k, n) in enumerate(SIZES):
        lim = np.sqrt(6.0 / k)
        W = rng.uniform(-lim, lim, (k, n))
        if l == 3:
            W[:, ERR + 1:] = 0.0
            W[:, :N_RAD] *= 0.5
        params.append([W, np.zeros(n)])
    return params


def forward(params, X, emulate_bf16=True):
    """X [B, 64] normalised features -> (z4 [B, 64], cache)."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    h = np.asarray(X, np.float64)
    cache = []
    for l, (W, b) in enumerate(params):
        a = r(h)
        z = a @ r(W) + b
        cache.append((a, z))
        h = np.maximum(z, 0.0) if l < 3 else z
    return h, cache


def loss_and_grad_out(z4, T):
    """MSE over the 48 radiance outputs plus the error head's regression onto
    log(per-sample MSE + eps) (target held constant).  Returns (loss, dL/dz4)."""
    B = z4.shape[0]
    diff = z4[:, :N_RAD] - T
    mse_s = (diff ** 2).mean(axis=1)
    e_t = np.log(mse_s + ERR_EPS)
    de = z4[:, ERR] - e_t
    loss = (diff ** 2).sum() / (B * N_RAD) + ERR_WEIGHT * (de ** 2).sum() / B
    g = np.zeros_like(z4)
    g[:, :N_RAD] = 2.0 * diff / (B * N_RAD)
    g[:, ERR] = ERR_WEIGHT * 2.0 * de / B
    return loss, g


def backward(params, cache, g4, emulate_bf16=True):
    """Gradients [(dW, db)] for dL/dz4 = g4, with the GPU's rounding points:
    the upstream gradient is rounded to bf16 before each product, bias
    gradients use the unrounded one."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    grads = [None] * 4
    g = g4
    for l in range(3, -1, -1):
        a, z = cache[l]
        G = r(g)
        grads[l] = [a.T @ G, g.sum(axis=0)]
        if l > 0:
            dh = G @ r(params[l][0]).T
            g = dh * (cache[l - 1][1] > 0)
    return grads


class Adam:
    """Adam with bias correction (Kingma & Ba 2015), FP64 state."""

    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps, self.t = lr, b1, b2, eps, 0
        self.m = [[np.zeros_like(p) for p in layer] for layer in params]
        self.v = [[np.zeros_like(p) for p in layer] for layer in params]

    def step(self, params, grads, lr=None):
        lr = self.lr if lr is None else lr
        self.t += 1
        c1 = 1.0 - self.b1 ** self.t
        c2 = 1.0 - self.b2 ** self.t
        for l in range(len(params)):
            for j in range(2):
                g = grads[l][j]
                self.m[l][j] = self.b1 * self.m[l][j] + (1 - self.b1) * g
                self.v[l][j] = self.b2 * self.v[l][j] + (1 - self.b2) * g * g
                params[l][j] = params[l][j] - lr * (self.m[l][j] / c1) / (np.sqrt(self.v[l][j] / c2) + self.eps)
        return params
