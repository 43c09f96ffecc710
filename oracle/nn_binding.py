"""ctypes binding of the surrogate-network restatement in oracle/liboracle.so
(nn_oracle.c).  Test infrastructure only; arrays are float64, NCHW."""
from __future__ import annotations

import ctypes

import numpy as np

from .binding import library as _library

_ready = False


def _lib() -> ctypes.CDLL:
    global _ready
    lib = _library()
    if not _ready:
        P, U, I, D, SZ = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
        lib.oracle_nn_conv_shape.argtypes = [I, ctypes.POINTER(U), ctypes.POINTER(U)]
        lib.oracle_nn_param_count.restype = SZ
        lib.oracle_conv3x3.argtypes = [P, P, P, U, U, U, U, U, I, P]
        lib.oracle_conv3x3_backward.argtypes = [P, P, P, U, U, U, U, U, P, P, P]
        lib.oracle_pool2.argtypes = [P, U, U, U, U, P]
        lib.oracle_upadd.argtypes = [P, P, U, U, U, U, P]
        lib.oracle_nn_forward.argtypes = [P, P, U, U, U, P]
        lib.oracle_nn_forward.restype = I
        lib.oracle_nn_loss_grad.argtypes = [P, P, P, U, U, U, P]
        lib.oracle_nn_loss_grad.restype = D
        lib.oracle_adam.argtypes = [P, P, P, P, SZ, D, D, D, D, U]
        _ready = True
    return lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def conv_shapes() -> list[tuple[int, int]]:
    lib = _lib()
    out = []
    for k in range(14):
        ci, co = ctypes.c_uint(), ctypes.c_uint()
        lib.oracle_nn_conv_shape(k, ctypes.byref(ci), ctypes.byref(co))
        out.append((ci.value, co.value))
    return out


def param_count() -> int:
    return int(_lib().oracle_nn_param_count())


def conv3x3(x, w, bias, relu=False):
    x, w = _f64(x), _f64(w)
    B, Cin, H, W = x.shape
    Cout = w.shape[0]
    y = np.empty((B, Cout, H, W))
    _lib().oracle_conv3x3(_p(x), _p(w), _p(None if bias is None else _f64(bias)), B, Cin, Cout, H, W,
                          1 if relu else 0, _p(y))
    return y


def conv3x3_backward(x, w, g, need_dx=True):
    x, w, g = _f64(x), _f64(w), _f64(g)
    B, Cin, H, W = x.shape
    Cout = w.shape[0]
    dx = np.empty_like(x) if need_dx else None
    dw, db = np.empty_like(w), np.empty(Cout)
    _lib().oracle_conv3x3_backward(_p(x), _p(w), _p(g), B, Cin, Cout, H, W, _p(dx), _p(dw), _p(db))
    return dx, dw, db


def pool2(x):
    x = _f64(x)
    B, C, H, W = x.shape
    y = np.empty((B, C, H // 2, W // 2))
    _lib().oracle_pool2(_p(x), B, C, H, W, _p(y))
    return y


def upadd(lo, skip):
    lo, skip = _f64(lo), _f64(skip)
    B, C, H, W = skip.shape
    z = np.empty_like(skip)
    _lib().oracle_upadd(_p(lo), _p(skip), B, C, H, W, _p(z))
    return z


def forward(params, x):
    params, x = _f64(params), _f64(x)
    B, _, H, W = x.shape
    out = np.empty((B, 3, H, W))
    if _lib().oracle_nn_forward(_p(params), _p(x), B, H, W, _p(out)) != 0:
        raise ValueError("extents must be positive multiples of 16")
    return out


def loss_grad(params, x, target):
    params, x, target = _f64(params), _f64(x), _f64(target)
    B, _, H, W = x.shape
    grad = np.empty_like(params)
    loss = _lib().oracle_nn_loss_grad(_p(params), _p(x), _p(target), B, H, W, _p(grad))
    return loss, grad


def adam(p, m, v, g, lr, b1, b2, eps, t):
    """In place on float64 arrays p, m, v."""
    assert p.dtype == m.dtype == v.dtype == np.float64
    g = _f64(g)
    _lib().oracle_adam(_p(p), _p(m), _p(v), _p(g), p.size, lr, b1, b2, eps, t)
