"""Device terminal-jump expectation of the control variate
(myql_debug_cv_jump_expectation: the closed form for Kou jumps and the
normal sums for Merton jumps, summed directly and read from the estimator's
log-spot table) against oracle/cv_jump.py, which integrates the down side
numerically; its derivative in log S against central differences."""
from __future__ import annotations

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, kou, make_model, merton
from oracle import cv_jump as cj

pytestmark = pytest.mark.gpu

STRIKES = np.array([80.0, 100.0, 120.5])
SPOTS = np.array([60.0, 93.0, 100.0, 118.0, 150.0])


@pytest.mark.parametrize("T", [0.1, 1.0])
@pytest.mark.parametrize("payoff", [capi.PAYOFF_VANILLA, capi.PAYOFF_CASH, capi.PAYOFF_ASSET])
@pytest.mark.parametrize("call", [True, False])
@pytest.mark.parametrize("jk", ["kou", "merton"])
def test_device_terminal_expectation_matches_oracle(engine, T, payoff, call, jk):
    jkind, jp = kou(1.0, 0.4, 10.0, 5.0) if jk == "kou" else merton(0.5, -0.1, 0.15)
    prm = jp if jk == "kou" else jp[:3]
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    ins = InstrumentSpec(payoff=payoff, option_type=capi.CALL if call else capi.PUT, strikes=STRIKES, maturity=T)
    print(f"cv jump expectation {jk} T={T} payoff={payoff} call={call}", flush=True)
    out = engine.cv_jump_expectation(m, ins, capi.GREEK_NONE, 100.0, SPOTS)
    assert out is not None
    tab, dr, ddr = out
    for i, S in enumerate(SPOTS):
        for k, K in enumerate(STRIKES):
            ref = cj.terminal_expectation(payoff, call, K, S, T, jk, prm)
            scale = 1.0 if payoff == capi.PAYOFF_CASH else K
            assert abs(dr[i, k] - ref) <= 1e-10 * scale, (S, K, dr[i, k], ref)
            assert abs(tab[i, k] - ref) <= 2e-7 * scale, (S, K, tab[i, k], ref)
            if i not in (1, 3) or k != 1:
                continue
            h = 1e-4
            up = cj.terminal_expectation(payoff, call, K, S * np.exp(h), T, jk, prm, include_zero=False)
            dn = cj.terminal_expectation(payoff, call, K, S * np.exp(-h), T, jk, prm, include_zero=False)
            assert abs(ddr[i, k] - (up - dn) / (2 * h)) <= 2e-6 * scale, (S, K, ddr[i, k], (up - dn) / (2 * h))


def test_uncovered_counts_and_wide_bandwidths_go_without_the_term(engine):
    # 40 counts per side cannot cover Poisson(60)
    jkind, jp = kou(100.0, 0.6, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    ins = InstrumentSpec(strikes=STRIKES, maturity=1.0)
    assert engine.cv_jump_expectation(m, ins, capi.GREEK_NONE, 100.0, SPOTS) is None
    # a smoothing window reaching below zero spot (K - 10 eps <= 0)
    jkind, jp = kou(1.0, 0.4, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    ins = InstrumentSpec(payoff=capi.PAYOFF_CASH, strikes=STRIKES, maturity=1.0, payoff_eps=9.0)
    assert engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, SPOTS) is None


@pytest.mark.parametrize("payoff", [capi.PAYOFF_CASH, capi.PAYOFF_ASSET])
@pytest.mark.parametrize("call", [True, False])
@pytest.mark.parametrize("jk", ["kou", "merton"])
def test_device_smoothed_digital_expectation(engine, payoff, call, jk):
    """Auto-smoothed digitals (Phi(+-(S - K) / eps)): the device integrates the
    smoothed payoff over its transition window against the jump density and
    takes the rest from the sharp tails. Checked against Monte Carlo of the
    smoothed payoff (5 SE), against the sharp oracle as eps -> 0, and its
    table and derivative against its own direct sum."""
    jkind, jp = kou(1.0, 0.4, 10.0, 5.0) if jk == "kou" else merton(0.5, -0.1, 0.15)
    prm = jp if jk == "kou" else jp[:3]
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    T, spots = 0.5, np.array([88.0, 100.0, 113.0])
    print(f"cv smoothed jump expectation {jk} payoff={payoff} call={call}", flush=True)
    for eps in (2.0, 0.5):
        ins = InstrumentSpec(payoff=payoff, option_type=capi.CALL if call else capi.PUT, strikes=STRIKES, maturity=T,
                             payoff_eps=eps)
        out = engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, spots)
        assert out is not None
        tab, dr, ddr = out
        scale = 1.0 if payoff == capi.PAYOFF_CASH else STRIKES[None, :]
        assert np.all(np.abs(tab - dr) <= 2e-7 * scale)
        for i, S in enumerate(spots):
            for k, K in enumerate(STRIKES):
                mean, se = cj.mc_terminal_expectation(payoff, call, K, S, T, jk, prm, n=1_000_000, seed=5 + i,
                                                      eps=eps)
                assert abs(dr[i, k] - mean) < 5 * se + 1e-12, (eps, S, K, dr[i, k], mean, se)
        # derivative of the n >= 1 part: central differences of the direct sum
        # (the no-jump term's own derivative removed)
        h = 1e-4
        up = engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, spots * np.exp(h))[1]
        dn = engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, spots * np.exp(-h))[1]
        lam = prm[0]
        m0 = -lam * T * (cj.kou_kappa(*prm[1:]) if jk == "kou" else (np.exp(prm[1] + 0.5 * prm[2] ** 2) - 1.0))
        z0 = lambda s: np.exp(-lam * T) * cj._payoff(payoff, call, s[:, None] * np.exp(m0), STRIKES[None, :], eps)
        fd = ((up - z0(spots * np.exp(h))) - (dn - z0(spots * np.exp(-h)))) / (2 * h)
        assert np.all(np.abs(ddr - fd) <= 1e-6 * scale + 1e-5 * np.abs(fd))
    # eps -> 0: the sharp expectation (strikes away from the spots' no-jump
    # atoms, so only the jumps' density matters)
    ins = InstrumentSpec(payoff=payoff, option_type=capi.CALL if call else capi.PUT, strikes=STRIKES, maturity=T,
                         payoff_eps=1e-3)
    dr = engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, spots)[1]
    for i, S in enumerate(spots):
        for k, K in enumerate(STRIKES):
            ref = cj.terminal_expectation(payoff, call, K, S, T, jk, prm)
            assert abs(dr[i, k] - ref) <= 2e-6 * (1.0 if payoff == capi.PAYOFF_CASH else K), (S, K, dr[i, k], ref)
