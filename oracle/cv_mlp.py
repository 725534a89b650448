"""ORACLE (test infrastructure only): FP64 numpy restatement of the learned
control-variate network of the engine (myquantlib_amd/csrc/cv_kernel.hpp),
which has no counterpart in the reference: MLP 5 -> 64 -> 64 -> 64 -> 64
with SiLU activations and a linear 64-output head (one output per strike
of the strip), parameters packed as W1[64][5], b1, W2[64][64], b2,
W3[64][64], b3, W4[64][64], b4[64] (12,864 floats).

  forward(p, f)        outputs for rows f (n x 5): (n x 64)
  tangents(p, f)       d out_k / d f_i for i = 1, 2, 3: (n x 64 x 3), by
                       forward-mode differentiation
  param_grad(p, f, t)  gradient of the mean over rows and outputs of
                       (out - t)^2 w.r.t. the parameters (t: n x 64)

The analytic derivatives are pinned by central finite differences in
tests/test_cv_reference.py.
"""
from __future__ import annotations

import numpy as np

H, NIN, NOUT = 64, 5, 64
_LAYOUT = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
           ("W4", (NOUT, H)), ("b4", (NOUT,))]
OFF = {}
_o = 0
for _k, _sh in _LAYOUT:
    OFF[_k] = (_o, _sh)
    _o += int(np.prod(_sh))
NPARAM = _o  # 12,864


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
    y = h3 @ P["W4"].T + P["b4"]
    return y, (f, h1, d1, h2, d2, h3, d3)


def forward(p, f):
    return _forward(unpack(p), f)[0]


def tangents(p, f):
    P = unpack(p)
    _, (_, _, d1, _, d2, _, d3) = _forward(P, f)
    out = np.zeros((np.asarray(f).shape[0], NOUT, 3))
    for i in range(3):
        t = d1 * P["W1"][:, 1 + i]
        t = (t @ P["W2"].T) * d2
        t = (t @ P["W3"].T) * d3
        out[:, :, i] = t @ P["W4"].T
    return out


def param_grad(p, f, t):
    P = unpack(p)
    y, (f, h1, d1, h2, d2, h3, d3) = _forward(P, f)
    t = np.asarray(t, dtype=np.float64)
    dy = 2.0 * (y - t) / y.size
    g3 = (dy @ P["W4"]) * d3
    g2 = (g3 @ P["W3"]) * d2
    g1 = (g2 @ P["W2"]) * d1
    G = {"W1": g1.T @ f, "b1": g1.sum(0), "W2": g2.T @ h1, "b2": g2.sum(0), "W3": g3.T @ h2,
         "b3": g3.sum(0), "W4": dy.T @ h3, "b4": dy.sum(0)}
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
