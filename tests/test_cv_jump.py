"""The control variate's terminal-jump expectation E[P(S e^J)] (oracle/cv_jump.py)
pinned by Monte Carlo with the jumps drawn as the reference draws them
(JumpPolicies.hpp:55-135: a Poisson count for the maturity, then each jump)
and by parity (vanilla call - put = S - K, cash call + put = 1, asset call +
put = S, since the compensation makes E[e^J] = 1)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cv_jump as cj

KOU = ("kou", (1.0, 0.4, 10.0, 5.0))
MERTON = ("merton", (0.5, -0.1, 0.15))


@pytest.mark.parametrize("jk", [KOU, MERTON], ids=["kou", "merton"])
@pytest.mark.parametrize("T", [0.1, 1.0])
def test_parity(jk, T):
    kind, prm = jk
    for K in (80.0, 100.0, 120.5):
        v = {(pf, c): cj.terminal_expectation(pf, c, K, 100.0, T, kind, prm) for pf in (0, 1, 2) for c in (True, False)}
        assert abs(v[(0, True)] - v[(0, False)] - (100.0 - K)) < 1e-9
        assert abs(v[(1, True)] + v[(1, False)] - 1.0) < 1e-11
        assert abs(v[(2, True)] + v[(2, False)] - 100.0) < 1e-9


@pytest.mark.parametrize("jk", [KOU, MERTON], ids=["kou", "merton"])
@pytest.mark.parametrize("payoff,call,K,S,T", [
    (0, True, 120.5, 100.0, 0.1), (0, True, 100.0, 97.0, 1.0), (0, False, 85.0, 104.0, 0.5),
    (1, True, 110.0, 100.0, 1.0), (1, False, 90.0, 100.0, 0.1), (2, True, 105.0, 100.0, 0.5)])
def test_matches_monte_carlo(jk, payoff, call, K, S, T):
    kind, prm = jk
    full = cj.terminal_expectation(payoff, call, K, S, T, kind, prm)
    part = cj.terminal_expectation(payoff, call, K, S, T, kind, prm, include_zero=False)
    mean, se = cj.mc_terminal_expectation(payoff, call, K, S, T, kind, prm, n=2_000_000, seed=11)
    assert abs(full - mean) < 5.0 * se + 1e-12, (full, mean, se)
    # the no-jump term is the payoff at the compensated spot, weight e^{-lambda T}
    lam = prm[0]
    m0 = -lam * T * (cj.kou_kappa(*prm[1:]) if kind == "kou" else (np.exp(prm[1] + 0.5 * prm[2] ** 2) - 1.0))
    assert abs(full - part - np.exp(-lam * T) * float(cj._payoff(payoff, call, S * np.exp(m0), K))) < 1e-12 * max(1.0, K)


def test_monte_carlo_check_is_sensitive():
    # a wrong compensation (kappa of the down side dropped) moves the value by
    # many standard errors: the Monte Carlo check would catch it
    kind, prm = KOU
    lam, p, e1, e2 = prm
    good = cj.terminal_expectation(0, True, 100.0, 100.0, 1.0, kind, prm)
    mean, se = cj.mc_terminal_expectation(0, True, 100.0, 100.0, 1.0, kind, prm, n=2_000_000, seed=12)
    bad = good * (1.0 + lam * (1.0 - p) * (e2 / (e2 + 1.0) - 1.0))
    assert abs(good - mean) < 5 * se and abs(bad - mean) > 20 * se
