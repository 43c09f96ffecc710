"""Warm-start surrogate on the GPU (include/lbm_surrogate.h) against the FP64
restatement in oracle/nn_oracle.c: forward, loss, every parameter gradient,
the Adam update, and training that lowers the loss deterministically."""
import numpy as np
import pytest

import portlbm_amd as pa
from oracle import nn_binding as nn
from portlbm_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def _case(seed, B, H, W, porosity=0.7):
    r = np.random.default_rng(seed)
    params = []
    for ci, co in nn.conv_shapes():
        params.append(r.normal(size=co * ci * 9) * np.sqrt(2.0 / (9 * ci)))
        params.append(r.normal(size=co) * 0.1)
    params = np.concatenate(params).astype(np.float32)
    x = r.normal(size=(B, 4, H, W)).astype(np.float32)
    x[:, 0] = (r.random(size=(B, H, W)) < porosity).astype(np.float32)
    t = r.normal(size=(B, 3, H, W)).astype(np.float32)
    return params, x, t


def _blocks(vec):
    off = 0
    for ci, co in nn.conv_shapes():
        yield vec[off:off + co * ci * 9]
        off += co * ci * 9
        yield vec[off:off + co]
        off += co


def _normwise(a, ref):
    return float(np.max(np.abs(a - ref)) / max(np.max(np.abs(ref)), 1e-30))


def test_param_count_matches_oracle():
    assert sg.param_count() == nn.param_count() == 201955


@pytest.mark.parametrize("B,H,W", [(2, 32, 48), (3, 16, 16), (1, 256, 256)])
def test_forward_matches_oracle(B, H, W):
    params, x, _ = _case(11 + H, B, H, W)
    with sg.Surrogate(H, W, max_batch=2) as net:  # B=3 runs in two chunks
        net.set_params(params)
        out = net.predict(x)
    ref = nn.forward(params.astype(np.float64), x.astype(np.float64))
    assert _normwise(out, ref) <= 1e-5


@pytest.mark.parametrize("B,H,W", [(2, 32, 48), (2, 64, 64)])
def test_loss_and_gradients_match_oracle(B, H, W):
    params, x, t = _case(21 + H, B, H, W)
    with sg.Surrogate(H, W, max_batch=B) as net:
        net.set_params(params)
        loss, g = net.loss_grad(x, t)
    rloss, rg = nn.loss_grad(params.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(rloss, rel=1e-5)
    for k, (a, r) in enumerate(zip(_blocks(g), _blocks(rg))):
        assert np.linalg.norm(a - r) <= 1e-4 * np.linalg.norm(r) + 1e-12, k


def test_adam_step_matches_oracle():
    params, x, t = _case(31, 2, 32, 32)
    lr = 1e-3
    with sg.Surrogate(32, 32, max_batch=2) as net:
        net.set_params(params)
        _, g0 = net.loss_grad(x, t)  # loss_grad leaves parameters and Adam state alone
        net.train_step(x, t, lr)
        p1 = net.params()
        _, g1 = net.loss_grad(x, t)
        net.train_step(x, t, lr)
        p2 = net.params()
    # the oracle's Adam driven by the very gradients the GPU stepped with
    p, m, v = params.astype(np.float64), np.zeros(params.size), np.zeros(params.size)
    nn.adam(p, m, v, g0.astype(np.float64), lr, 0.9, 0.999, 1e-8, 1)
    np.testing.assert_allclose(p1, p, rtol=0, atol=1e-6 * lr + 2e-7 * np.max(np.abs(p)))
    nn.adam(p, m, v, g1.astype(np.float64), lr, 0.9, 0.999, 1e-8, 2)
    np.testing.assert_allclose(p2, p, rtol=0, atol=1e-5 * lr + 4e-7 * np.max(np.abs(p)))
    assert np.max(np.abs(p2 - params)) > 0.5 * lr  # the steps did move the parameters


def _smooth_task(n, H, W, seed):
    r = np.random.default_rng(seed)
    x = np.zeros((n, 4, H, W), dtype=np.float32)
    t = np.zeros((n, 3, H, W), dtype=np.float32)
    yy, xx = np.mgrid[0:H, 0:W] / np.array([H, W])[:, None, None]
    for i in range(n):
        a, b = r.uniform(-1, 1, size=2)
        mask = (r.random((H, W)) < 0.8).astype(np.float32)
        x[i, 0], x[i, 1], x[i, 2], x[i, 3] = mask, 2 * xx - 1, a, b
        t[i, 0] = mask * (1 - xx) * (1 + 0.3 * a)
        t[i, 1] = mask * np.sin(np.pi * yy) * (1 + 0.5 * b)
        t[i, 2] = mask * 0.2 * np.cos(2 * np.pi * xx) * a
    return x, t


def test_training_lowers_loss_and_is_deterministic():
    x, t = _smooth_task(24, 32, 32, 5)
    runs = []
    for _ in range(2):
        with sg.Surrogate(32, 32, max_batch=8) as net:
            net.init_params(3)
            net.dataset_reserve(24)
            net.dataset_add(x[:16], t[:16])
            net.dataset_add(x[16:], t[16:])
            assert net.dataset_size() == 24
            h1 = net.train(150, 300, 8, 2e-3, 1e-4, seed=9)
            h2 = net.train(150, 300, 8, 2e-3, 1e-4, seed=9)  # continues the schedule
            runs.append((np.concatenate([h1, h2]), net.params()))
    hist, p = runs[0]
    assert np.all(np.isfinite(hist))
    assert np.mean(hist[-20:]) < 0.1 * np.mean(hist[:5])
    np.testing.assert_array_equal(runs[1][0], hist)
    np.testing.assert_array_equal(runs[1][1], p)


def test_rejects_bad_arguments():
    with pytest.raises(pa.LbmError):
        sg.Surrogate(30, 32)
    with sg.Surrogate(16, 16, max_batch=2) as net:
        with pytest.raises(pa.LbmError):  # parameters never set
            net.predict(np.zeros((1, 4, 16, 16), np.float32))
        net.init_params(1)
        with pytest.raises(ValueError):
            net.predict(np.zeros((1, 4, 16, 32), np.float32))
        with pytest.raises(pa.LbmError):  # batch above max_batch
            net.loss_grad(np.ones((3, 4, 16, 16), np.float32), np.zeros((3, 3, 16, 16), np.float32))
        with pytest.raises(pa.LbmError):  # no fluid node
            net.loss_grad(np.zeros((1, 4, 16, 16), np.float32), np.zeros((1, 3, 16, 16), np.float32))
        net.dataset_reserve(2)
        with pytest.raises(pa.LbmError):
            net.dataset_add(np.ones((3, 4, 16, 16), np.float32), np.zeros((3, 3, 16, 16), np.float32))


def _bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _torch_bf16_network(params, x):
    """The network with each convolution's input and weights rounded to BF16 and
    exact (FP64) accumulation -- what the matrix-core inference computes up to
    its FP32 accumulation."""
    import torch
    import torch.nn.functional as F
    ws, off = [], 0
    for ci, co in nn.conv_shapes():
        w = torch.from_numpy(_bf16(params[off:off + co * ci * 9]).reshape(co, ci, 3, 3))
        off += co * ci * 9
        ws.append((w, torch.from_numpy(params[off:off + co].astype(np.float64))))
        off += co

    def c(k, t, relu=True):
        y = F.conv2d(torch.from_numpy(_bf16(t.numpy())), ws[k][0], ws[k][1], padding=1)
        return F.relu(y) if relu else y
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")  # noqa: E731
    xt = torch.from_numpy(x.astype(np.float64))
    a2 = c(1, c(0, xt))
    a4 = c(3, c(2, F.avg_pool2d(a2, 2)))
    a6 = c(5, c(4, F.avg_pool2d(a4, 2)))
    a8 = c(7, c(6, F.avg_pool2d(a6, 2)))
    a9 = c(8, F.avg_pool2d(a8, 2))
    a10 = c(9, up(a9) + a8)
    a11 = c(10, up(a10) + a6)
    a12 = c(11, up(a11) + a4)
    a13 = c(12, up(a12) + a2)
    return c(13, a13, relu=False).numpy()


@pytest.mark.parametrize("B,H,W", [(2, 32, 48), (1, 64, 80)])
def test_bf16_matrix_core_inference(B, H, W):
    params, x, _ = _case(41 + H, B, H, W)
    with sg.Surrogate(H, W, max_batch=B) as net:
        net.set_params(params)
        out = net.predict(x, bf16=True)
        ref32 = net.predict(x)
    emul = _torch_bf16_network(params, x)
    # against the BF16 emulation: only accumulation order and FP32 rounding differ
    assert _normwise(out, emul) <= 2e-4
    # against FP32 inference: the BF16 rounding of 14 layers' inputs and weights
    assert _normwise(out, ref32) <= 3e-2
