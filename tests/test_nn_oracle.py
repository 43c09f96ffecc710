"""Pins the FP64 restatement of the warm-start surrogate (oracle/nn_oracle.c):
forward operators against torch's CPU conv2d/avg_pool2d/interpolate, every
backward against central finite differences of the oracle's own forward, and
Adam against a hand-evaluated step and the closed form of its first step."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import nn_binding as nn


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_architecture_shapes_and_count():
    shapes = nn.conv_shapes()
    assert shapes == [(4, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64), (64, 64),
                      (64, 48), (48, 32), (32, 16), (16, 16), (16, 3)]
    assert nn.param_count() == sum(9 * ci * co + co for ci, co in shapes) == 201955


@pytest.mark.parametrize("relu", [False, True])
def test_conv3x3_matches_torch(relu):
    r = _rng(1)
    x, w, b = r.normal(size=(2, 3, 7, 9)), r.normal(size=(5, 3, 3, 3)), r.normal(size=5)
    ref = F.conv2d(torch.from_numpy(x), torch.from_numpy(w), torch.from_numpy(b), padding=1)
    if relu:
        ref = F.relu(ref)
    np.testing.assert_allclose(nn.conv3x3(x, w, b, relu), ref.numpy(), rtol=1e-12, atol=1e-12)


def test_pool_and_upadd_match_torch():
    r = _rng(2)
    x = r.normal(size=(2, 3, 8, 6))
    np.testing.assert_allclose(nn.pool2(x), F.avg_pool2d(torch.from_numpy(x), 2).numpy(), rtol=1e-14)
    lo, skip = r.normal(size=(2, 3, 4, 3)), r.normal(size=(2, 3, 8, 6))
    ref = F.interpolate(torch.from_numpy(lo), scale_factor=2, mode="nearest") + torch.from_numpy(skip)
    np.testing.assert_array_equal(nn.upadd(lo, skip), ref.numpy())


def test_conv_backward_finite_differences():
    r = _rng(3)
    x, w, g = r.normal(size=(2, 3, 5, 6)), r.normal(size=(4, 3, 3, 3)), r.normal(size=(2, 4, 5, 6))
    b = r.normal(size=4)
    dx, dw, db = nn.conv3x3_backward(x, w, g)
    # L = <g, conv(x)>  =>  dL/dx = dx, dL/dw = dw, dL/db = db
    L = lambda xx, ww, bb: float(np.sum(g * nn.conv3x3(xx, ww, bb)))
    h = 1e-6
    for arr, grad, name in ((x, dx, "x"), (w, dw, "w"), (b, db, "b")):
        flat = arr.reshape(-1)
        for i in r.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[i]
            flat[i] = old + h
            lp = L(x, w, b)
            flat[i] = old - h
            lm = L(x, w, b)
            flat[i] = old
            assert abs((lp - lm) / (2 * h) - grad.reshape(-1)[i]) <= 1e-7 * max(1.0, abs(grad.reshape(-1)[i])), (name, i)


def _net_case(seed, B=2, H=16, W=32, porosity=0.7):
    r = _rng(seed)
    shapes = nn.conv_shapes()
    params = []
    for ci, co in shapes:
        params.append(r.normal(size=co * ci * 9) * np.sqrt(2.0 / (9 * ci)))
        params.append(r.normal(size=co) * 0.1)
    params = np.concatenate(params)
    x = r.normal(size=(B, 4, H, W))
    x[:, 0] = (r.random(size=(B, H, W)) < porosity).astype(np.float64)
    t = r.normal(size=(B, 3, H, W))
    return params, x, t


def test_network_forward_matches_torch_composition():
    params, x, _ = _net_case(4)
    shapes = nn.conv_shapes()
    ws, off = [], 0
    for ci, co in shapes:
        w = torch.from_numpy(params[off:off + co * ci * 9].reshape(co, ci, 3, 3))
        off += co * ci * 9
        ws.append((w, torch.from_numpy(params[off:off + co])))
        off += co
    c = lambda k, t, relu=True: (F.relu if relu else (lambda v: v))(F.conv2d(t, ws[k][0], ws[k][1], padding=1))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    xt = torch.from_numpy(x)
    a2 = c(1, c(0, xt))
    a4 = c(3, c(2, F.avg_pool2d(a2, 2)))
    a6 = c(5, c(4, F.avg_pool2d(a4, 2)))
    a8 = c(7, c(6, F.avg_pool2d(a6, 2)))
    a9 = c(8, F.avg_pool2d(a8, 2))
    a10 = c(9, up(a9) + a8)
    a11 = c(10, up(a10) + a6)
    a12 = c(11, up(a11) + a4)
    a13 = c(12, up(a12) + a2)
    out = c(13, a13, relu=False)
    np.testing.assert_allclose(nn.forward(params, x), out.numpy(), rtol=1e-11, atol=1e-11)


def test_network_gradient_finite_differences():
    params, x, t = _net_case(5)
    loss, grad = nn.loss_grad(params, x, t)
    m = x[:, 0:1]
    out = nn.forward(params, x)
    assert loss == pytest.approx(float(np.sum(m * (out - t) ** 2) / (3 * m.sum())), rel=1e-13)
    r = _rng(6)
    # one parameter from each weight block and each bias block, plus random ones
    shapes = nn.conv_shapes()
    picks, off = [], 0
    for ci, co in shapes:
        picks.append(off + int(r.integers(co * ci * 9)))
        off += co * ci * 9
        picks.append(off + int(r.integers(co)))
        off += co
    picks += list(r.choice(params.size, size=8, replace=False))
    # A difference quotient straddling a ReLU kink is off by O(h); such picks
    # are rare, so all must agree to 1e-4 and nearly all to 1e-6.
    h = 1e-6
    rel = []
    for i in picks:
        old = params[i]
        params[i] = old + h
        lp = nn.loss_grad(params, x, t)[0]
        params[i] = old - h
        lm = nn.loss_grad(params, x, t)[0]
        params[i] = old
        fd = (lp - lm) / (2 * h)
        rel.append(abs(fd - grad[i]) / max(abs(grad[i]), 1e-3))
    rel = np.array(rel)
    assert rel.max() <= 1e-4, rel
    assert np.mean(rel <= 1e-6) >= 0.9, rel


def test_network_rejects_bad_extents():
    params, x, t = _net_case(7, H=16, W=32)
    with pytest.raises(ValueError):
        nn.forward(params, x[:, :, :, :24])
    assert np.isnan(nn.loss_grad(params, x[:, :, :8], t[:, :, :8])[0])


def test_adam_known_step():
    p, m, v = np.array([1.0, -2.0, 0.5]), np.zeros(3), np.zeros(3)
    g = np.array([0.1, -0.3, 0.0])
    nn.adam(p, m, v, g, 0.01, 0.9, 0.999, 1e-8, 1)
    # first bias-corrected step: m_hat = g, v_hat = g^2  =>  p -= lr * g / (|g| + eps)
    np.testing.assert_allclose(p, [1.0 - 0.01 * 0.1 / (0.1 + 1e-8), -2.0 + 0.01 * 0.3 / (0.3 + 1e-8), 0.5],
                               rtol=1e-15)
    g2 = np.array([0.2, 0.1, -0.4])
    nn.adam(p, m, v, g2, 0.01, 0.9, 0.999, 1e-8, 2)
    m2 = 0.9 * (0.1 * g) + 0.1 * g2
    v2 = 0.999 * (0.001 * g * g) + 0.001 * g2 * g2
    np.testing.assert_allclose(m, m2, rtol=1e-14)
    np.testing.assert_allclose(v, v2, rtol=1e-14)
