"""TEST INFRASTRUCTURE ONLY: float64 numpy model of the neural radiance cache.

The radiance cache has no counterpart in the reference; this module restates
the network the HIP kernels implement (pyre_amd/csrc/nrc.hip) so the tests
can check the device forward pass, loss, gradients and Adam update. The
gradients here are themselves pinned by central finite differences in
tests/test_nrc_oracle.py.

Network: 13 raw features per query -> 128-wide encoding -> 5 hidden layers
of 64 with ReLU -> 16 outputs (the first 3 are RGB radiance, the rest carry
zero weight), no biases (the constant-1 encoding slot acts as the bias of
the first layer). Loss: relative L2, mean over queries and RGB of
(pred - target)^2 / (sg(pred)^2 + 0.01).
"""
from __future__ import annotations

import numpy as np

N_FREQ = 12
D_IN = 108          # 72 position frequencies + 16 direction SH + 16 normal SH + roughness + albedo(3)
K_IN = 128          # padded input width (slot 108 = constant 1, then zeros)
HIDDEN = 64
N_HIDDEN = 5
N_OUT = 16          # padded output width; outputs 0..2 are RGB
LAYERS = [(K_IN, HIDDEN)] + [(HIDDEN, HIDDEN)] * (N_HIDDEN - 1) + [(HIDDEN, N_OUT)]
REL_EPS = 0.01


def n_params() -> int:
    return sum(a * b for a, b in LAYERS)


def sh4(d):
    """Real spherical harmonics l = 0..3 (16 values) of unit vectors d (n, 3)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    return np.stack([
        np.full_like(x, 0.28209479177387814),
        -0.48860251190291987 * y,
        0.48860251190291987 * z,
        -0.48860251190291987 * x,
        1.0925484305920792 * x * y,
        -1.0925484305920792 * y * z,
        0.94617469575755997 * zz - 0.31539156525251999,
        -1.0925484305920792 * x * z,
        0.54627421529603959 * (xx - yy),
        0.59004358992664352 * y * (-3.0 * xx + yy),
        2.8906114426405538 * x * y * z,
        0.45704579946446572 * y * (1.0 - 5.0 * zz),
        0.3731763325901154 * z * (5.0 * zz - 3.0),
        0.45704579946446572 * x * (1.0 - 5.0 * zz),
        1.4453057213202769 * z * (xx - yy),
        0.59004358992664352 * x * (-xx + 3.0 * yy),
    ], axis=1)


def encode(feats, bmin, bmax):
    """feats (n, 13): position, outgoing direction, normal, roughness, albedo."""
    f = np.asarray(feats, np.float64)
    n = len(f)
    out = np.zeros((n, K_IN))
    p = (f[:, 0:3] - np.asarray(bmin, np.float64)) / (np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64))
    for a in range(3):
        for k in range(N_FREQ):
            arg = (2.0 ** k) * np.pi * p[:, a]
            out[:, a * 2 * N_FREQ + 2 * k] = np.sin(arg)
            out[:, a * 2 * N_FREQ + 2 * k + 1] = np.cos(arg)
    out[:, 72:88] = sh4(f[:, 3:6])
    out[:, 88:104] = sh4(f[:, 6:9])
    out[:, 104] = f[:, 9]
    out[:, 105:108] = f[:, 10:13]
    out[:, 108] = 1.0
    return out


def init_weights(seed: int):
    """He-uniform init from a numpy PCG64 stream (the device copies these)."""
    rng = np.random.default_rng(seed)
    ws = []
    for li, (a, b) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / a)
        w = rng.uniform(-lim, lim, (a, b))
        if li == len(LAYERS) - 1:
            w[:, 3:] = 0.0
        if li == 0:
            w[D_IN + 1:, :] = 0.0  # padding rows after the constant slot
        ws.append(w)
    return ws


def flatten(ws):
    return np.concatenate([w.reshape(-1) for w in ws])


def unflatten(v):
    out, o = [], 0
    for a, b in LAYERS:
        out.append(np.asarray(v[o:o + a * b], np.float64).reshape(a, b))
        o += a * b
    return out


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(ws, x, quant=False):
    """Returns (outputs (n, 16), list of layer inputs). quant=True mimics the
    device: bf16 weights and activations, fp32-class accumulation."""
    acts = [bf16(x) if quant else x]
    h = acts[0]
    for li, w in enumerate(ws):
        z = h @ (bf16(w) if quant else w)
        if li < len(ws) - 1:
            h = np.maximum(z, 0.0)
            if quant:
                h = bf16(h)
            acts.append(h)
        else:
            return z, acts
    raise AssertionError


def loss_and_grad_out(pred, target):
    """Relative L2 over RGB; returns (loss, dL/dpred (n, 16))."""
    p = pred[:, :3]
    t = np.asarray(target, np.float64)
    n = len(p)
    denom = p * p + REL_EPS
    loss = float(np.mean((p - t) ** 2 / denom))
    g = np.zeros_like(pred)
    g[:, :3] = 2.0 * (p - t) / denom / (3 * n)
    return loss, g


def backward(ws, acts, gout):
    grads = [None] * len(ws)
    dz = gout
    for li in range(len(ws) - 1, -1, -1):
        grads[li] = acts[li].T @ dz
        if li > 0:
            dz = (dz @ ws[li].T) * (acts[li] > 0)
    grads[-1][:, 3:] = 0.0
    return grads


def loss_fn(ws, x, target):
    pred, _ = forward(ws, x)
    return loss_and_grad_out(pred, target)[0]


def gradients(ws, x, target):
    pred, acts = forward(ws, x)
    loss, g = loss_and_grad_out(pred, target)
    return loss, backward(ws, acts, g)


def adam(params, grads, m, v, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15):
    """One Adam step (Kingma & Ba) on flat arrays; t is the 1-based step."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
