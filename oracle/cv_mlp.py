"""ORACLE (test infrastructure only): FP64 numpy restatement of the learned
control-variate network of the engine (myquantlib_amd/csrc/cv_kernel.hpp),
which has no counterpart in the reference: MLP 6 -> 64 -> 64 -> 64 -> 1 with
SiLU activations, parameters packed as W1[64][6], b1, W2[64][64], b2,
W3[64][64], b3, w4[64], b4 (8,833 floats).

  forward(p, f)        network outputs for rows f (n x 6)
  input_grad(p, f)     d net / d f (n x 6), by reverse-mode differentiation
  param_grad(p, f, t)  gradient of mean (net - t)^2 w.r.t. the parameters

The analytic gradients are pinned by central finite differences in
tests/test_cv_reference.py.
"""
from __future__ import annotations

import numpy as np

H, NIN, NPARAM = 64, 6, 8833
OFF = {"W1": (0, (H, NIN)), "b1": (384, (H,)), "W2": (448, (H, H)), "b2": (4544, (H,)),
       "W3": (4608, (H, H)), "b3": (8704, (H,)), "w4": (8768, (H,)), "b4": (8832, (1,))}


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {k: p[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in OFF.items()}


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s + z * s * (1.0 - s)


def _forward(P, f):
    f = np.asarray(f, dtype=np.float64)
    a1 = f @ P["W1"].T + P["b1"]
    h1, d1 = _silu(a1)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2, d2 = _silu(a2)
    a3 = h2 @ P["W3"].T + P["b3"]
    h3, d3 = _silu(a3)
    y = h3 @ P["w4"] + P["b4"][0]
    return y, (f, h1, d1, h2, d2, h3, d3)


def forward(p, f):
    return _forward(unpack(p), f)[0]


def input_grad(p, f):
    P = unpack(p)
    _, (_, _, d1, _, d2, _, d3) = _forward(P, f)
    g3 = d3 * P["w4"]
    g2 = (g3 @ P["W3"]) * d2
    g1 = (g2 @ P["W2"]) * d1
    return g1 @ P["W1"]


def param_grad(p, f, t):
    P = unpack(p)
    y, (f, h1, d1, h2, d2, h3, d3) = _forward(P, f)
    n = f.shape[0]
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n
    g3 = (dy[:, None] * P["w4"]) * d3
    g2 = (g3 @ P["W3"]) * d2
    g1 = (g2 @ P["W2"]) * d1
    G = {"W1": g1.T @ f, "b1": g1.sum(0), "W2": g2.T @ h1, "b2": g2.sum(0), "W3": g3.T @ h2,
         "b3": g3.sum(0), "w4": dy @ h3, "b4": np.array([dy.sum()])}
    out = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        out[o:o + int(np.prod(sh))] = G[k].reshape(-1)
    return out, float(np.mean((y - t) ** 2))


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        fan_in = sh[1] if len(sh) == 2 else H
        p[o:o + int(np.prod(sh))] = rng.normal(0.0, scale / np.sqrt(fan_in), int(np.prod(sh)))
    return p
