"""CPU oracle of the learned cons2prim guess network -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatement of the engine's 8-64-64-2 GELU network: inputs,
forward pass, mean-squared-error loss and its gradient, in the flat parameter
layout W1 [8][64], b1, W2 [64][64], b2, W3 [64][2], b3. Pinned by central
finite differences (tests/test_mlp_oracle.py). The network has no reference
counterpart; the reference starts its root find from the previous primitives
(Project/CPU/Src/srmhd.cc:438-442).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 8, 64, 2
SIZES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)
# GELU as the engine evaluates it: x Phi~(x) with Phi~(x) = 1/2 + t P(t^2),
# t = clip(x, -c, c), P of degree 4 (the FP32 constants of mgpu_mlp.hpp, here
# in float64); within 1.22e-3 of the exact x Phi(x) over |x| <= 40.
GELU_C = float(np.float32(3.38253573))
GELU_P = [float(np.float32(v)) for v in (3.93638938e-01, -5.91898904e-02, 6.53467361e-03, -3.95613815e-04,
                                         9.83280758e-06)]


def unpack(p):
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SIZES])


def features(cons8):
    """(D, Sx, Sy, Sz, tau, Bx, By, Bz) -> the 8 network inputs."""
    c = np.asarray(cons8, dtype=np.float64)
    D, tau = c[:, 0], c[:, 4]
    E = tau + D
    return np.stack([np.log(D) * 0.2, c[:, 1] / E, c[:, 2] / E, c[:, 3] / E, np.log(E / D) * 0.25,
                     c[:, 5] / np.sqrt(E), c[:, 6] / np.sqrt(E), c[:, 7] / np.sqrt(E)], axis=1)


def targets(W, p, D, tau):
    return np.stack([np.log(W), np.log(p / (tau + D))], axis=1)


def _poly(t2):
    p = np.zeros_like(t2)
    for c in reversed(GELU_P):
        p = p * t2 + c
    return p


def _dpoly(t2):
    p = np.zeros_like(t2)
    for k in reversed(range(1, len(GELU_P))):
        p = p * t2 + k * GELU_P[k]
    return p


def gelu(x):
    x = np.asarray(x, dtype=np.float64)
    t = np.clip(x, -GELU_C, GELU_C)
    return x * (0.5 + t * _poly(t * t))


def gelu_grad(x):
    x = np.asarray(x, dtype=np.float64)
    t = np.clip(x, -GELU_C, GELU_C)
    t2 = t * t
    P = _poly(t2)
    dphi = np.where(np.abs(x) < GELU_C, P + 2.0 * t2 * _dpoly(t2), 0.0)
    return 0.5 + t * P + x * dphi


def forward(p, x):
    w = unpack(p)
    a1 = x @ w["W1"] + w["b1"]
    h1 = gelu(a1)
    a2 = h1 @ w["W2"] + w["b2"]
    h2 = gelu(a2)
    y = h2 @ w["W3"] + w["b3"]
    return y, (a1, h1, a2, h2)


def loss_and_grad(p, x, t):
    """Mean over samples and outputs of (y - t)^2, and its gradient."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    y, (a1, h1, a2, h2) = forward(p, x)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dy = 2.0 * (y - t) / n
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    da2 = (dy @ w["W3"].T) * gelu_grad(a2)
    g["W2"], g["b2"] = h1.T @ da2, da2.sum(0)
    da1 = (da2 @ w["W2"].T) * gelu_grad(a1)
    g["W1"], g["b1"] = x.T @ da1, da1.sum(0)
    return loss, pack(g)
