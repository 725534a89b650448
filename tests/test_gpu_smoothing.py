"""GPU parity of the automatic digital-smoothing bandwidth
(MonteCarloPricer.hpp:31-74, Payoffs.hpp:39-52).

* The engine's bandwidth equals the one the reference's own variance helpers
  give (tests/golden/smoothing_ref.json, compiled from ModelParams.hpp:54-80):
  on identical draws a digital priced with payoff_eps < 0 equals, bit for
  bit, the same digital with the fixture's bandwidth given -- ten models x
  four maturities, scalar strike and strip.
* Auto-smoothed cash and asset strips in Essential and Full mode equal the
  CPU oracle on identical draws (FP64 paths: 1e-9 on prices and first-order
  Greeks, 1e-6 on second-order ones, as in test_gpu_parity.py).
"""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, mc_config, kou, merton

from test_gpu_parity import ESS, FULL, compare
from test_smoothing import FIELDS, load, price_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("strip", [True, False])
def test_engine_auto_bandwidth_equals_reference_helpers(engine, strip):
    fx = load()
    for c in fx["cases"]:
        auto, given = price_pair(engine, fx, c, strip, capi.GREEK_ESSENTIAL)
        for f in FIELDS:
            assert np.array_equal(auto[f], given[f]), (c["model"], c["T"], f, auto[f], given[f])


def test_engine_bandwidth_test_is_sensitive(engine):
    fx = load()
    c = next(x for x in fx["cases"] if x["model"] == "bates_kou" and x["T"] == 1.0)
    auto, given = price_pair(engine, fx, c, True, capi.GREEK_ESSENTIAL, scale=1.0 + 1e-6)
    assert not np.array_equal(auto["price"], given["price"])


BATES_KOU = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=kou(1.0, 0.4, 10.0, 5.0)[0],
                       jump=kou(1.0, 0.4, 10.0, 5.0)[1])
BATES = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=merton(0.5, -0.1, 0.15)[0],
                   jump=merton(0.5, -0.1, 0.15)[1])


@pytest.mark.parametrize("mode", [capi.GREEK_ESSENTIAL, capi.GREEK_FULL])
@pytest.mark.parametrize("payoff", [capi.PAYOFF_CASH, capi.PAYOFF_ASSET])
@pytest.mark.parametrize("which", ["bates_kou_nv", "bates_nci"])
def test_auto_smoothed_digital_strip_matches_oracle(engine, oracle, mode, payoff, which):
    m, sch = (BATES_KOU, capi.SCHEME_NV) if which == "bates_kou_nv" else (BATES, capi.SCHEME_NCI)
    inst = InstrumentSpec(payoff=payoff, strikes=np.linspace(70.0, 130.0, 16), maturity=0.5)
    cfg = mc_config(1 << 14, 50, seed=42, num_streams=512)
    g = engine.mc(m, [sch], inst, mode, cfg, 100.0, 0.02, 0.0)
    o = oracle.mc(m, [sch], inst, mode, cfg, 100.0, 0.02, 0.0)
    compare(g, o, FULL if mode == capi.GREEK_FULL else ESS, 1e-9)
    # smoothing is on: the prices differ from the sharp (GreekMode::None) ones
    sharp = engine.mc(m, [sch], inst, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)
    assert not np.allclose(sharp["price"], g["price"], rtol=1e-6)
