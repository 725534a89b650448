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


def test_smoothed_digitals_and_uncovered_counts_go_without_the_term(engine):
    jkind, jp = kou(1.0, 0.4, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    ins = InstrumentSpec(payoff=capi.PAYOFF_CASH, strikes=STRIKES, maturity=1.0)
    assert engine.cv_jump_expectation(m, ins, capi.GREEK_ESSENTIAL, 100.0, SPOTS) is None
    # 40 counts per side cannot cover Poisson(60)
    jkind, jp = kou(100.0, 0.6, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jkind, jump=jp)
    ins = InstrumentSpec(strikes=STRIKES, maturity=1.0)
    assert engine.cv_jump_expectation(m, ins, capi.GREEK_NONE, 100.0, SPOTS) is None
