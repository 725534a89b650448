"""The C++20 drop-in layer on the GPU: tests/cpp/dropin_check (reference-
style user code over include/myql, built by __graft_entry__.build()) runs
each case through MonteCarloPricer / MultiLevelMonteCarloPricer; the same
case through the Python binding of the C-ABI must give identical numbers,
and SchemeNV outside its domain must raise the reference's exception text.
"""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import (Heston, InstrumentSpec, make_model, mc_config, merton,
                                    mlmc_config)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_check")
S0, R, Q = 100.0, 0.02, 0.0
HESTON = Heston(2.0, 0.04, 0.3, -0.7, 0.04)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cases():
    assert os.path.exists(BIN), "tests/cpp/dropin_check is built by __graft_entry__.build()"
    print("running tests/cpp/dropin_check", flush=True)
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return {d["case"]: d for d in (json.loads(line) for line in out.stdout.splitlines())}


def same(a, b):
    np.testing.assert_array_equal(np.atleast_1d(np.asarray(a, dtype=float)),
                                  np.atleast_1d(np.asarray(b, dtype=float)))


def test_heston_nv_call(cases, engine):
    got = cases["heston_nv_call"]
    res = engine.mc(make_model(1, [HESTON]), [capi.SCHEME_NV], InstrumentSpec(strikes=[100.0]),
                    capi.GREEK_ESSENTIAL, mc_config(1 << 16, 50), S0, R, Q)
    for f in ("price", "price_std_err", "delta", "gamma"):
        same(got[f], res[f])
    assert got["total_steps"] == res["total_steps"]


def test_bates_nci_strip_full(cases, engine):
    got = cases["bates_nci_strip"]
    jk, jp = merton(0.5, -0.1, 0.15)
    model = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)
    res = engine.mc(model, [capi.SCHEME_NCI], InstrumentSpec(strikes=[90.0, 100.0, 110.0], maturity=0.5),
                    capi.GREEK_FULL, mc_config(1 << 14, 25), S0, R, Q)
    for f in ("price", "delta", "gamma", "theta", "rho", "charm"):
        same(got[f], res[f])
    same(got["vega0"], res["vega"][0])
    same(got["vanna0"], res["vanna"][0])


def test_barrier_up_and_out(cases, engine):
    got = cases["heston_uo_call"]
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0,
                          barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)
    res = engine.mc(make_model(1, [HESTON]), [capi.SCHEME_NV], inst, capi.GREEK_NONE,
                    mc_config(1 << 15, 64), S0, R, Q)
    same(got["price"], res["price"])
    same(got["price_std_err"], res["price_std_err"])


def test_arith_asian(cases, engine):
    got = cases["heston_asian_call"]
    inst = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], fixed_strike=True)
    res = engine.mc(make_model(1, [HESTON]), [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                    mc_config(1 << 15, 32), S0, R, Q)
    same(got["price"], res["price"])
    same(got["delta"], res["delta"])


def test_mlmc(cases, engine):
    got = cases["heston_mlmc_call"]
    cfg = mlmc_config(epsilon=0.05, L_min=2, L_max=4, N0=4096, base_steps=2)
    res = engine.mlmc(make_model(1, [HESTON]), [capi.SCHEME_NV], InstrumentSpec(strikes=[100.0]),
                      capi.GREEK_NONE, cfg, S0, R, Q)
    same(got["price"], res["price"])
    same(got["price_std_err"], res["price_std_err"])
    assert got["levels"] == res["n_levels"]


def test_nv_feller_raises_reference_message(cases):
    assert cases["nv_feller"]["error"] == "SchemeNV violation: sigma^2 > 4*kappa*theta"


def test_fourier_pricer(cases, engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    model = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)
    got = cases["bates_fourier_strip"]
    res = engine.fourier(model, InstrumentSpec(strikes=[90.0, 100.0, 110.0], maturity=0.5), capi.GREEK_FULL,
                         S0, R, Q)
    for f in ("price", "delta", "theta"):
        same(got[f], res[f])
    same(got["vega0"], res["vega"][0])
    got = cases["bates_fourier_digital"]
    res = engine.fourier(model, InstrumentSpec(strikes=[95.0], maturity=1.0, payoff=capi.PAYOFF_CASH,
                                               option_type=capi.PUT), capi.GREEK_ESSENTIAL, S0, R, Q)
    for f in ("price", "delta", "gamma"):
        same(got[f], res[f])
