"""Multilevel Monte Carlo against plain Monte Carlo at the finest level
(the reference's benchmarks/mlmc_verifier.cpp:153-402 compares the two).

With L_min = L_max = L the MLMC estimator's expectation telescopes to the
expectation of the plain estimator with base_steps * 2^L steps:
E[P_0] + sum_l E[P_f,l - P_c,l] = E[P_L], because the coarse path of level
l is distributed as the fine path of level l - 1 (CoupledASVJStepper.hpp:
225-349). So the two estimates, on independent streams, must agree within
3 combined standard errors with no bias allowance -- for a European under
Bates (NV, Merton jumps at maturity) and under Bates with SchemeNCI (the
coarse variance copies the fine one), an arithmetic Asian, the up-and-out
call of the MLMC workload (FP64 and FP32 paths) and a floating lookback put
under Bates-Kou (per-step jumps: the coarse jump is the sum of the fine
ones).
"""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, mc_config, merton, kou, mlmc_config

pytestmark = pytest.mark.gpu

S0, R, Q = 100.0, 0.05, 0.02
H = Heston(2.0, 0.04, 0.4, -0.7, 0.04)
H1 = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
H_HI = Heston(2.0, 0.04, 1.0, -0.7, 0.04)


def _bates(h, lam=0.1, mu=-0.05, dl=0.15):
    jk, jp = merton(lam, mu, dl)
    return make_model(1, [h], jump_kind=jk, jump=jp)


def _bates_kou(h):
    jk, jp = kou(0.1, 0.3, 10.0, 10.0)
    return make_model(1, [h], jump_kind=jk, jump=jp)


CASES = {
    "bates_european_nv": (_bates(H), capi.SCHEME_NV, InstrumentSpec(strikes=[100.0], maturity=1.0)),
    "bates_european_nci": (_bates(H_HI, lam=0.5, mu=-0.1), capi.SCHEME_NCI,
                           InstrumentSpec(strikes=[100.0], maturity=1.0)),
    "heston_arith_asian": (make_model(1, [H]), capi.SCHEME_NV,
                           InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=1.0)),
    "heston_up_and_out": (make_model(1, [H1]), capi.SCHEME_NV,
                          InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0, maturity=1.0,
                                         barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)),
    "bates_kou_floating_lookback_put": (_bates_kou(H), capi.SCHEME_NV,
                                        InstrumentSpec(kind=capi.INSTR_LOOKBACK, option_type=capi.PUT,
                                                       strikes=[100.0], maturity=1.0, fixed_strike=False)),
}


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mlmc_equals_plain_mc_at_finest_level(engine, name, prec):
    m, sch, inst = CASES[name]
    base, L = 4, 3
    steps = base << L
    plain = engine.mc(m, [sch], inst, capi.GREEK_NONE,
                      mc_config(1 << 21, steps, seed=5, precision=prec, num_streams=1 << 15), S0, R, Q)
    p, sp = float(plain["price"][0]), float(plain["price_std_err"][0])
    ml = engine.mlmc(m, [sch], inst, capi.GREEK_NONE,
                     mlmc_config(epsilon=2e-3 * p, L_min=L, L_max=L, N0=20000, base_steps=base, seed=42,
                                 precision=prec, num_streams=1 << 15), S0, R, Q)
    assert ml["n_levels"] == L + 1
    se = float(np.hypot(ml["price_std_err"], sp))
    z = (ml["price"] - p) / se
    assert abs(z) < 3.0, (name, ml["price"], p, ml["price_std_err"], sp, z)
    # the level corrections shrink: the coupled paths stay close
    md = np.abs(np.asarray(ml["mean_diffs"][1:L + 1]))
    assert md[-1] < 0.1 * p, md
