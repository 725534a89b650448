"""The FP64 reference of the control-variate network (oracle/cv_mlp.py) is
pinned by central finite differences: its forward-mode tangents of the 64
outputs and its parameter gradient of the squared loss must match the
numerical ones."""
import numpy as np

from oracle import cv_mlp


def test_tangents_match_finite_differences():
    p = cv_mlp.random_params(3)
    f = np.random.default_rng(4).normal(size=(5, 5))
    g = cv_mlp.tangents(p, f)
    h = 1e-6
    for i in range(3):
        e = np.zeros(5)
        e[1 + i] = h
        num = (cv_mlp.forward(p, f + e) - cv_mlp.forward(p, f - e)) / (2 * h)
        assert np.allclose(g[:, :, i], num, rtol=1e-6, atol=1e-9), i


def test_param_gradient_matches_finite_differences():
    p = cv_mlp.random_params(5)
    rng = np.random.default_rng(6)
    f = rng.normal(size=(7, 5))
    t = rng.normal(size=(7, 64))
    g, _ = cv_mlp.param_grad(p, f, t)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(sh))), 4, replace=False)
                          for o, sh in cv_mlp.OFF.values()])
    for i in idx:
        e = np.zeros(cv_mlp.NPARAM)
        e[i] = h
        lp = np.mean((cv_mlp.forward(p + e, f) - t) ** 2)
        lm = np.mean((cv_mlp.forward(p - e, f) - t) ** 2)
        assert abs(g[i] - (lp - lm) / (2 * h)) <= 1e-6 * max(1.0, abs(g[i])), i


def test_layout_is_the_engines():
    assert cv_mlp.NPARAM == 12864
    assert cv_mlp.OFF["W4"] == (8704, (64, 64)) and cv_mlp.OFF["b4"] == (12800, (64,))
