"""The automatic digital-smoothing bandwidth (MonteCarloPricer.hpp:31-74) in
the CPU oracle, pinned by the reference's own variance helpers
(models/asvj/data/ModelParams.hpp:54-80) compiled from its headers:
tests/golden/smoothing_ref.json holds the expected average variance and the
bandwidth of a scalar strike (K = 110, eps = sqrt(v) K sqrt(T) / 10) and of a
strike strip (S0 = 100 in place of K) for ten models at four maturities.

The bandwidth only enters through the smoothed payoff, so each case prices a
cash-or-nothing digital twice on identical draws -- once with payoff_eps < 0
(automatic) and once with the fixture's bandwidth given explicitly -- and the
two results must be equal bit for bit; a bandwidth off by one part in 1e6
must not be. GreekMode::None never smooths, whatever the bandwidth.
"""
import json
import os

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, mc_config

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "smoothing_ref.json")
STRIP = np.linspace(80.0, 120.0, 5)


def load():
    with open(FIXTURE) as f:
        return json.load(f)


def fixture_model(fx, name):
    m = fx["models"][name]
    hs = [Heston(**h) for h in m["heston"]]
    if m["jump"] is None:
        jk, jp = capi.JUMP_NONE, (0, 0, 0, 0)
    else:
        jk = capi.JUMP_MERTON if m["jump"]["kind"] == "merton" else capi.JUMP_KOU
        jp = tuple(m["jump"]["p"])
    # SchemeNV where sigma^2 <= 4 kappa theta, else SchemeExact
    schemes = [capi.SCHEME_NV if h.sigma ** 2 <= 4 * h.kappa * h.theta else capi.SCHEME_EXACT for h in hs]
    return make_model(m["n_factors"], hs, vol=m["vol"], jump_kind=jk, jump=jp), schemes


def price_pair(pricer, fx, case, strip, mode, payoff=capi.PAYOFF_CASH, paths=512, scale=1.0):
    """(automatic, explicit) results of one fixture case on identical draws."""
    m, schemes = fixture_model(fx, case["model"])
    T = case["T"]
    eps = case["eps_strip" if strip else "eps_scalar"] * scale
    K = STRIP if strip else [fx["K_scalar"]]
    cfg = mc_config(paths, max(2, int(round(20 * T))), seed=11)
    out = []
    for e in (-1.0, eps):
        inst = InstrumentSpec(payoff=payoff, strikes=K, maturity=T, is_strip=strip, payoff_eps=e)
        out.append(pricer.mc(m, schemes, inst, mode, cfg, fx["S0"], 0.02, 0.0))
    return out


FIELDS = ("price", "price_std_err", "delta", "gamma")


def test_fixture_is_consistent():
    """The fixture's bandwidths follow from its variances by the rule."""
    fx = load()
    assert len(fx["cases"]) == 40
    for c in fx["cases"]:
        s = np.sqrt(c["avg_var"]) * np.sqrt(c["T"]) / 10.0
        assert c["eps_strip"] == pytest.approx(s * fx["S0"], rel=1e-15)
        assert c["eps_scalar"] == pytest.approx(s * fx["K_scalar"], rel=1e-15)


@pytest.mark.parametrize("strip", [True, False])
def test_oracle_auto_bandwidth_equals_reference_helpers(oracle, strip):
    fx = load()
    for c in fx["cases"]:
        auto, given = price_pair(oracle, fx, c, strip, capi.GREEK_ESSENTIAL)
        for f in FIELDS:
            assert np.array_equal(auto[f], given[f]), (c["model"], c["T"], f, auto[f], given[f])


def test_oracle_bandwidth_test_is_sensitive(oracle):
    """A bandwidth one part in 1e6 away changes the smoothed prices."""
    fx = load()
    c = next(x for x in fx["cases"] if x["model"] == "bates_kou" and x["T"] == 1.0)
    auto, given = price_pair(oracle, fx, c, True, capi.GREEK_ESSENTIAL, scale=1.0 + 1e-6)
    assert not np.array_equal(auto["price"], given["price"])


def test_oracle_none_mode_is_sharp(oracle):
    """GreekMode::None prices the sharp indicator whatever the bandwidth
    (smoothing applies in Essential and Full only, MonteCarloPricer.hpp:51-53)."""
    fx = load()
    c = next(x for x in fx["cases"] if x["model"] == "heston" and x["T"] == 0.5)
    auto, given = price_pair(oracle, fx, c, True, capi.GREEK_NONE, scale=3.0)
    assert np.array_equal(auto["price"], given["price"])
