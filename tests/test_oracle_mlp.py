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


def test_adam_step1_closed_form():
    """Step 1 of Adam with bias correction is p - lr * g / (|g| + eps) for every
    element (m/c1 = g, v/c2 = g^2): a missing or mis-powered bias correction
    moves p by a different amount."""
    rng = np.random.default_rng(3)
    params = [[rng.standard_normal((4, 3)), rng.standard_normal(3)]]
    p0 = [[a.copy() for a in layer] for layer in params]
    grads = [[rng.standard_normal((4, 3)), rng.standard_normal(3) * 1e-3]]
    lr, eps = 2e-3, 1e-8
    mlp.Adam(params, lr=lr, eps=eps).step(params, grads)
    for j in range(2):
        g = grads[0][j]
        np.testing.assert_allclose(params[0][j], p0[0][j] - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)


def test_adam_matches_torch_optim():
    """Six steps with changing gradients and a changing learning rate agree
    with torch.optim.Adam (a library, FP64) to 1e-12 relative."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    shapes = [(5, 4), (4,), (4, 2), (2,)]
    params = [[rng.standard_normal(shapes[0]), rng.standard_normal(shapes[1])],
              [rng.standard_normal(shapes[2]), rng.standard_normal(shapes[3])]]
    tp = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for layer in params for a in layer]
    opt_t = torch.optim.Adam(tp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    opt_o = mlp.Adam(params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8)
    for step in range(6):
        lr = 1e-3 * (0.7 ** step)
        grads = [[rng.standard_normal(shapes[0]) * (step + 1), rng.standard_normal(shapes[1])],
                 [rng.standard_normal(shapes[2]), rng.standard_normal(shapes[3]) * 1e-4]]
        for t, g in zip(tp, [g for layer in grads for g in layer]):
            t.grad = torch.tensor(g, dtype=torch.float64)
        for group in opt_t.param_groups:
            group["lr"] = lr
        opt_t.step()
        opt_o.step(params, grads, lr=lr)
        for t, a in zip(tp, [a for layer in params for a in layer]):
            ref = t.detach().numpy()
            assert np.max(np.abs(a - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1.0)
