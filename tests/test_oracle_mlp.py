"""Pins for the FP64 MLP oracle: finite differences and torch autograd (a
library, not the reference) agree with its backward; the bf16-emulating
path stays close to the exact one."""
import numpy as np
import pytest

from oracle import mlp


def _data(B=64, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(B, mlp.IN)), rng.normal(size=(B, mlp.N_RAD))


def _loss(params, X, T, emulate=False):
    z4, _ = mlp.forward(params, X, emulate)
    return mlp.loss_and_grad_out(z4, T)[0]


def test_backward_matches_finite_differences(monkeypatch):
    # the error head's target is held constant in the backward (a stop-gradient),
    # which finite differences cannot see: check the radiance loss alone
    monkeypatch.setattr(mlp, "ERR_WEIGHT", 0.0)
    params = mlp.init_params(3)
    X, T = _data()
    z4, cache = mlp.forward(params, X, emulate_bf16=False)
    _, g4 = mlp.loss_and_grad_out(z4, T)
    grads = mlp.backward(params, cache, g4, emulate_bf16=False)
    rng = np.random.default_rng(0)
    h = 1e-6
    for l in range(4):
        for j in range(2):
            P = params[l][j]
            for _ in range(6):
                idx = tuple(rng.integers(s) for s in P.shape)
                if l == 3 and (idx[-1] >= mlp.ERR):
                    continue  # error head (weight 0 here) and padding: no loss
                old = P[idx]
                P[idx] = old + h
                lp = _loss(params, X, T)
                P[idx] = old - h
                lm = _loss(params, X, T)
                P[idx] = old
                fd = (lp - lm) / (2 * h)
                an = grads[l][j][idx]
                assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (l, j, idx, fd, an)


def test_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    params = mlp.init_params(5)
    X, T = _data(32, 2)
    z4, cache = mlp.forward(params, X, emulate_bf16=False)
    _, g4 = mlp.loss_and_grad_out(z4, T)
    grads = mlp.backward(params, cache, g4, emulate_bf16=False)
    tp = [[torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)] for W, b in params]
    h = torch.tensor(X)
    for l, (W, b) in enumerate(tp):
        h = h @ W + b
        if l < 3:
            h = torch.relu(h)
    diff = h[:, :mlp.N_RAD] - torch.tensor(T)
    mse_s = (diff ** 2).mean(dim=1)
    e_t = torch.log(mse_s + mlp.ERR_EPS).detach()
    loss = (diff ** 2).sum() / (32 * mlp.N_RAD) + mlp.ERR_WEIGHT * ((h[:, mlp.ERR] - e_t) ** 2).sum() / 32
    loss.backward()
    for l in range(4):
        assert np.allclose(tp[l][0].grad.numpy(), grads[l][0], rtol=1e-10, atol=1e-12)
        assert np.allclose(tp[l][1].grad.numpy(), grads[l][1], rtol=1e-10, atol=1e-12)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0, 0.0, 1e-3])
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # 1 + 2^-8 is a tie -> even
    assert r[2] == 1.0078125               # above the tie -> up
    assert r[3] == -3.0 and r[4] == 0.0
    assert abs(r[5] - 1e-3) / 1e-3 < 2 ** -8


def test_emulated_path_close_to_exact():
    params = mlp.init_params(7)
    X, T = _data(128, 4)
    z_e, c_e = mlp.forward(params, X, True)
    z_x, c_x = mlp.forward(params, X, False)
    assert np.abs(z_e - z_x).max() < 0.05 * np.abs(z_x).max()
    _, g_e = mlp.loss_and_grad_out(z_e, T)
    gr_e = mlp.backward(params, c_e, g_e, True)
    _, g_x = mlp.loss_and_grad_out(z_x, T)
    gr_x = mlp.backward(params, c_x, g_x, False)
    for l in range(4):
        num = np.linalg.norm(gr_e[l][0] - gr_x[l][0])
        assert num <= 0.05 * np.linalg.norm(gr_x[l][0])


def test_adam_decreases_loss():
    params = mlp.init_params(1)
    X, T = _data(256, 9)
    T = T * 0.1
    opt = mlp.Adam(params, lr=3e-3)
    l0 = _loss(params, X, T)
    for _ in range(60):
        z4, cache = mlp.forward(params, X, False)
        _, g4 = mlp.loss_and_grad_out(z4, T)
        opt.step(params, mlp.backward(params, cache, g4, False))
    assert _loss(params, X, T) < 0.5 * l0
