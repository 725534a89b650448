"""Weights of the multilevel Richardson-Romberg estimator (ML2R, Lemaire &
Pages 2017) in the oracle, pinned by closed forms and by the Vandermonde
conditions they are defined by. The reference has no ML2R (README.md:332-335
lists it as future work); the engine's weights (a Lagrange closed form) are
checked against these on the GPU through the MLMC result's level_weights."""
import numpy as np
import pytest


def test_richardson_romberg_two_levels(oracle):
    # depth 2 on refiners 1, 2: 2^a/(2^a - 1) P_fine - 1/(2^a - 1) P_coarse,
    # i.e. level means weighted 1 and 2^a/(2^a - 1)
    for a in (0.5, 1.0, 2.0):
        W = oracle.mlmc_level_weights(1, a)
        assert np.allclose(W, [1.0, 2.0 ** a / (2.0 ** a - 1.0)], rtol=0, atol=1e-14)


def test_three_levels_closed_form(oracle):
    # alpha = 1, nodes x = 1, 1/2, 1/4: w = (1/3, -2, 8/3), W = (1, 2/3, 8/3)
    W = oracle.mlmc_level_weights(2, 1.0)
    assert np.allclose(W, [1.0, 2.0 / 3.0, 8.0 / 3.0], rtol=0, atol=1e-13)
    # alpha = 2 (the reference's default weak order), x = 1, 1/4, 1/16:
    # w = (1/45, -4/9, 64/45), W = (1, 44/45, 64/45)
    W = oracle.mlmc_level_weights(2, 2.0)
    assert np.allclose(W, [1.0, 44.0 / 45.0, 64.0 / 45.0], rtol=0, atol=1e-13)


@pytest.mark.parametrize("alpha", [0.5, 1.0, 1.5, 2.0])
@pytest.mark.parametrize("L", [0, 1, 2, 3, 5, 7])
def test_weights_cancel_bias_terms(oracle, alpha, L):
    """w_l = W_l - W_{l+1} satisfy sum w = 1 and sum w x^k = 0, k = 1..L,
    x_l = 2^(-alpha l): an estimate whose level-l mean is P(h 2^-l) minus
    P(h 2^-(l-1)) loses the terms h^alpha .. h^(alpha L) of its bias."""
    W = oracle.mlmc_level_weights(L, alpha)
    w = W - np.append(W[1:], 0.0)
    x = 2.0 ** (-alpha * np.arange(L + 1))
    assert abs(w.sum() - 1.0) < 1e-12
    for k in range(1, L + 1):
        assert abs(np.sum(w * x ** k)) < 1e-10 * np.sum(np.abs(w))
    # a bias expansion c_1 h^a + ... + c_L h^(aL) is removed exactly
    rng = np.random.default_rng(L)
    c = rng.normal(size=L + 1)
    P = lambda h: 3.0 + sum(c[k] * h ** (alpha * k) for k in range(1, L + 1))  # noqa: E731
    h0 = 0.5
    means = [P(h0)] + [P(h0 * 2.0 ** -l) - P(h0 * 2.0 ** -(l - 1)) for l in range(1, L + 1)]
    assert abs(np.dot(W, means) - 3.0) < 1e-10


def test_giles_weights_are_one(oracle):
    assert np.array_equal(oracle.mlmc_level_weights(4, 1.0, ml2r=False), np.ones(5))
