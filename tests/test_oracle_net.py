"""Pins of the network oracle (oracle/net_oracle.c, the checker of the
learned preconditioner's GPU kernels): its gradient against central finite
differences of its own loss with exact operands, and the BF16 rounding
against exactly representable values and ties to even."""
import ctypes as C

import numpy as np

import oracle_py


def lib():
    L = oracle_py.lib()
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.on_num_params.restype = C.c_long
    L.on_forward.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, C.c_int]
    L.on_loss_grad.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, C.POINTER(C.c_float), fp, C.c_int]
    return L


def random_problem(B=2, H=16, W=8, seed=0):
    L = lib()
    rng = np.random.default_rng(seed)
    n = L.on_num_params()
    p = (0.2 * rng.standard_normal(n)).astype(np.float32)
    x = rng.uniform(-1, 1, (B, H, W, 2)).astype(np.float32)
    t = rng.uniform(-1, 1, (B, H, W)).astype(np.float32)
    return p, x, t


def loss_grad(p, x, t, round_=0):
    L = lib()
    B, H, W, _ = x.shape
    g = np.zeros_like(p)
    loss = C.c_float()
    assert L.on_loss_grad(B, H, W, p, np.ascontiguousarray(x), np.ascontiguousarray(t), C.byref(loss), g, round_) == 0
    return loss.value, g


def test_gradient_matches_finite_differences():
    p, x, t = random_problem()
    _, g = loss_grad(p, x, t)
    rng = np.random.default_rng(1)
    # layers 0 and 6 only use 2 input / 1 output channels: test live
    # parameters spread over every layer and the biases
    idx = rng.choice(np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max()), 40, replace=False)
    errs = []
    for q in idx:
        h = 1e-2 * max(abs(p[q]), 0.05)
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (loss_grad(pp, x, t)[0] - loss_grad(pm, x, t)[0]) / (pp[q] - pm[q])
        errs.append(abs(fd - g[q]) / max(abs(g[q]), 1e-6))
    errs = np.array(errs)
    # ReLU kinks may spoil a few differences; the bulk must agree closely
    assert np.median(errs) < 2e-3, errs
    assert np.mean(errs < 2e-2) >= 0.9, errs


def test_bf16_rounding_changes_the_loss_slightly():
    p, x, t = random_problem(seed=3)
    l0, g0 = loss_grad(p, x, t, 0)
    l1, g1 = loss_grad(p, x, t, 1)
    assert l1 != l0 and abs(l1 - l0) < 1e-2 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)


def test_forward_is_the_loss_forward():
    L = lib()
    p, x, t = random_problem(B=1, seed=5)
    y = np.zeros(t.shape, np.float32)
    assert L.on_forward(1, 16, 8, p, x, y, 0) == 0
    loss, _ = loss_grad(p, x, t)
    assert abs(loss - np.sum((y - t) ** 2) / np.sum(t ** 2)) < 1e-5 * loss
