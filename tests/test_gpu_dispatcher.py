"""The runtime dispatcher on the GPU: the drop-in's own
include/myql/dispatcher (tests/cpp/dispatch_check_dropin) and, where the
reference's headers are present, the reference's myql/dispatcher/*.hpp
compiled unchanged over the drop-in's other headers
(oracle/_ref/dispatch_check). Both are built by __graft_entry__.build() from
tests/cpp/dispatch_check.cpp (build_model, build_instrument, build_european,
price_mc, price_fourier, DispatchResult). Every case must give exactly the
numbers the Python binding of the same C-ABI gives for the same call;
SchemeNV outside its domain must surface the reference's exception text;
the drop-in sets DispatchResult::time_ms, which the reference leaves 0.
"""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, kou, make_model, mc_config, merton

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINS = {"reference": os.path.join(ROOT, "oracle", "_ref", "dispatch_check"),
        "dropin": os.path.join(ROOT, "tests", "cpp", "dispatch_check_dropin")}
S0, R, Q = 100.0, 0.02, 0.0
H1 = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
H2 = Heston(1.0, 0.03, 0.25, -0.5, 0.02)
H_BATES = Heston(2.0, 0.04, 1.0, -0.7, 0.04)
STRIP = dict(strikes=[90.0, 100.0, 110.0], maturity=0.5)

pytestmark = pytest.mark.gpu


def _run(which):
    b = BINS[which]
    if not os.path.exists(b):
        pytest.skip(f"{os.path.relpath(b, ROOT)} not built (the reference's dispatcher is built only where its "
                    "headers are present)")
    print(f"running {os.path.relpath(b, ROOT)}", flush=True)
    out = subprocess.run([b], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return {d["case"]: d for d in (json.loads(line) for line in out.stdout.splitlines())}


@pytest.fixture(scope="module", params=["dropin", "reference"])
def cases(request):
    return _run(request.param)


def test_dropin_sets_time_ms():
    got = _run("dropin")
    for name, d in got.items():
        if "error" not in d and name != "scheme_nv_violation":
            assert d["time_ms"] > 0.0, name


def same(a, b):
    np.testing.assert_array_equal(np.atleast_1d(np.asarray(a, dtype=float)),
                                  np.atleast_1d(np.asarray(b, dtype=float)))


def check_mc(got, res, full):
    """DispatchResult fields against the binding's result."""
    pairs = [("prices", "price"), ("prices_std_err", "price_std_err"), ("deltas", "delta"),
             ("gammas", "gamma"), ("deltas_std_err", "delta_std_err"), ("gammas_std_err", "gamma_std_err")]
    if full:
        pairs += [("thetas", "theta"), ("rhos", "rho"), ("thetas_std_err", "theta_std_err"),
                  ("rhos_std_err", "rho_std_err")]
    for d, f in pairs:
        same(got[d], res[f])
    if full:
        for k in range(2):
            same(got[f"vegas{k}"], res["vega"][k])
            same(got[f"vegas_std_err{k}"], res["vega_std_err"][k])


def test_bates_nci_vanilla_strip_full(cases, engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    res = engine.mc(make_model(1, [H_BATES], jump_kind=jk, jump=jp), [capi.SCHEME_NCI], InstrumentSpec(**STRIP),
                    capi.GREEK_FULL, mc_config(1 << 14, 25), S0, R, Q)
    check_mc(cases["bates_nci_vanilla_call_full"], res, True)


def test_heston_up_and_out(cases, engine):
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0,
                          barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)
    res = engine.mc(make_model(1, [H1]), [capi.SCHEME_NV], inst, capi.GREEK_NONE, mc_config(1 << 15, 64), S0, R, Q)
    got = cases["heston_nv_up_and_out_call"]
    same(got["prices"], res["price"])
    same(got["prices_std_err"], res["price_std_err"])
    assert got["deltas"] == [] and "vegas0" not in got      # GreekMode::None: prices only


def test_double_heston_arith_asian(cases, engine):
    inst = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], fixed_strike=True)
    res = engine.mc(make_model(2, [H1, H2]), [capi.SCHEME_NV, capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                    mc_config(1 << 14, 32), S0, R, Q)
    check_mc(cases["double_heston_nv_arith_asian_call"], res, False)


def test_bates_kou_floating_lookback_put(cases, engine):
    jk, jp = kou(1.0, 0.4, 10.0, 5.0)
    inst = InstrumentSpec(kind=capi.INSTR_LOOKBACK, option_type=capi.PUT, fixed_strike=False, strikes=[0.0])
    res = engine.mc(make_model(1, [H1], jump_kind=jk, jump=jp), [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                    mc_config(1 << 14, 32), S0, R, Q)
    check_mc(cases["bates_kou_nv_lookback_floating_put"], res, False)


def test_merton_cash_put_strip(cases, engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    inst = InstrumentSpec(payoff=capi.PAYOFF_CASH, option_type=capi.PUT, **STRIP)
    res = engine.mc(make_model(0, vol=0.2, jump_kind=jk, jump=jp), [], inst, capi.GREEK_ESSENTIAL,
                    mc_config(1 << 14, 10), S0, R, Q)
    check_mc(cases["merton_cash_put_essential"], res, False)


def test_fourier_heston_cash_call_full(cases, engine):
    res = engine.fourier(make_model(1, [H1]), InstrumentSpec(payoff=capi.PAYOFF_CASH, **STRIP),
                         capi.GREEK_FULL, S0, R, Q)
    got = cases["fourier_heston_cash_call_full"]
    for d, f in [("prices", "price"), ("deltas", "delta"), ("gammas", "gamma"), ("thetas", "theta"),
                 ("rhos", "rho")]:
        same(got[d], res[f])
    for k in range(2):
        same(got[f"vegas{k}"], res["vega"][k])
    assert got["prices_std_err"] == []          # Fourier results carry no standard errors


def test_fourier_double_bates_kou_put(cases, engine):
    jk, jp = kou(1.0, 0.4, 10.0, 5.0)
    res = engine.fourier(make_model(2, [H1, H2], jump_kind=jk, jump=jp),
                         InstrumentSpec(option_type=capi.PUT, **STRIP), capi.GREEK_ESSENTIAL, S0, R, Q)
    got = cases["fourier_double_bates_kou_vanilla_put"]
    for d, f in [("prices", "price"), ("deltas", "delta"), ("gammas", "gamma")]:
        same(got[d], res[f])
    assert got["thetas"] == []


def test_scheme_nv_violation_message(cases):
    assert cases["scheme_nv_violation"].get("error") == "SchemeNV violation: sigma^2 > 4*kappa*theta"
