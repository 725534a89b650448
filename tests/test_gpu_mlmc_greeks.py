"""Multilevel Greeks and the multilevel Richardson-Romberg estimator (ML2R).

The reference computes each level path's spot-up and spot-down payoffs
(calculate_to_buffer, MultiLevelMonteCarloPricer.hpp:280,321-324) and drops
them; its README.md:332-335 lists MLMC Greeks and ML2R as future work. The
engine sums those buffers' central differences per level (delta and gamma)
and can weight the level means by the Richardson weights (ML2R). Checked:
 - identical draws: GPU = oracle (FP64), for both estimators, with the
   oracle's weights solved from the Vandermonde system independently of the
   engine's closed form;
 - telescoping: at a fixed finest level the multilevel delta and gamma have
   the plain estimator's expectation at that level (independent streams,
   within 3 combined standard errors);
 - against the Gil-Pelaez delta and gamma of a European call;
 - ML2R cancels the scheme's leading bias terms: at a coarse finest level
   it lands on the Fourier price where the plain multilevel sum is visibly
   biased.
"""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, mc_config, mlmc_config

pytestmark = pytest.mark.gpu

S0, R, Q = 100.0, 0.02, 0.0
H1 = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
EURO = InstrumentSpec(strikes=[100.0], maturity=1.0)
UO = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0, maturity=1.0,
                    barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)
ASIAN = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=1.0)


def close(a, b, se, tol=1e-9):
    return abs(a - b) <= tol * max(abs(b), se, 1e-300)


@pytest.mark.parametrize("estimator", [capi.MLMC_GILES, capi.MLMC_ML2R])
@pytest.mark.parametrize("inst", [EURO, UO], ids=["european", "up_and_out"])
def test_greeks_fp64_match_oracle(engine, oracle, inst, estimator):
    m = make_model(1, [H1])
    cfg = mlmc_config(epsilon=0.05, L_min=2, L_max=5, N0=2000, base_steps=4, num_streams=1024,
                      alpha=1.0 if estimator == capi.MLMC_ML2R else 2.0, estimator=estimator)
    print("mlmc greeks vs oracle:", inst.kind, estimator, flush=True)
    g = engine.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL, cfg, S0, R, Q)
    c = oracle.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL, cfg, S0, R, Q)
    assert g["n_levels"] == c["n_levels"]
    assert np.array_equal(g["paths_per_level"], c["paths_per_level"])
    assert np.allclose(g["level_weights"], c["level_weights"], rtol=1e-12, atol=1e-12)
    if estimator == capi.MLMC_GILES:
        assert np.array_equal(g["level_weights"], np.ones(g["n_levels"]))
    for f in ("price", "delta", "gamma"):
        assert close(g[f], c[f], c[f + "_std_err"]), (f, g[f], c[f])
        assert close(g[f + "_std_err"], c[f + "_std_err"], 0.0), (f, g[f + "_std_err"], c[f + "_std_err"])
    # the estimate is the weighted sum of the level means
    assert abs(np.dot(g["level_weights"], g["mean_diffs"]) - g["price"]) < 1e-12 * abs(g["price"]) + 1e-15


def test_none_mode_leaves_greeks_zero(engine):
    m = make_model(1, [H1])
    cfg = mlmc_config(epsilon=0.05, L_min=2, L_max=4, N0=2000, base_steps=4, num_streams=1024)
    g = engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_NONE, cfg, S0, R, Q)
    assert g["delta"] == 0.0 and g["gamma"] == 0.0 and g["delta_std_err"] == 0.0


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("inst", [UO, ASIAN], ids=["up_and_out", "arith_asian"])
def test_greeks_equal_plain_mc_at_finest_level(engine, inst, prec):
    """With L_min = L_max = L the level sums telescope to the plain
    estimator at base_steps * 2^L steps: same expected price, delta, gamma."""
    m = make_model(1, [H1])
    base, L = 4, 3
    plain = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                      mc_config(1 << 21, base << L, seed=5, precision=prec, num_streams=1 << 15), S0, R, Q)
    ml = engine.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                     mlmc_config(epsilon=2e-3 * float(plain["price"][0]), L_min=L, L_max=L, N0=40000,
                                 base_steps=base, seed=42, precision=prec, num_streams=1 << 15), S0, R, Q)
    print("finest-level", inst.kind, prec, {f: (ml[f], float(plain[f][0])) for f in ("price", "delta", "gamma")},
          flush=True)
    for f in ("price", "delta", "gamma"):
        se = float(np.hypot(ml[f + "_std_err"], float(plain[f + "_std_err"][0])))
        z = (ml[f] - float(plain[f][0])) / se
        assert abs(z) < 3.0, (f, ml[f], float(plain[f][0]), se, z)


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_delta_gamma_against_fourier(engine, prec):
    m = make_model(1, [H1])
    fo = engine.fourier(m, EURO, capi.GREEK_ESSENTIAL, S0, R, Q)
    p = float(fo["price"][0])
    ml = engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_ESSENTIAL,
                     mlmc_config(epsilon=5e-4 * p, L_min=2, L_max=8, N0=20000, base_steps=8, seed=42,
                                 precision=prec, num_streams=1 << 16), S0, R, Q)
    print("mlmc vs fourier", prec, {f: (ml[f], ml[f + "_std_err"], float(fo[f][0])) for f in ("price", "delta", "gamma")},
          flush=True)
    for f in ("price", "delta", "gamma"):
        z = (ml[f] - float(fo[f][0])) / ml[f + "_std_err"]
        assert abs(z) < 3.5, (f, ml[f], float(fo[f][0]), ml[f + "_std_err"], z)


def test_ml2r_removes_discretisation_bias(engine):
    """One step per year at level 0 and three levels (finest 4 steps): the
    plain multilevel sum estimates the 4-step price and carries the scheme's
    bias (~190 standard errors here); ML2R's weights for weak order 2 -- the
    reference's default alpha, and the NV scheme's order: successive level
    means shrink by ~4 -- cancel its h^2 and h^4 terms."""
    m = make_model(1, [H1])
    fo = float(engine.fourier(m, EURO, capi.GREEK_NONE, S0, R, Q)["price"][0])
    out = {}
    for est in (capi.MLMC_GILES, capi.MLMC_ML2R):
        out[est] = engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_NONE,
                               mlmc_config(epsilon=5e-5 * fo, L_min=2, L_max=2, N0=20000, base_steps=1, seed=42,
                                           alpha=2.0, estimator=est, num_streams=1 << 16), S0, R, Q)
    g, r = out[capi.MLMC_GILES], out[capi.MLMC_ML2R]
    print("bias: fourier", fo, "giles", g["price"], g["price_std_err"], "ml2r", r["price"], r["price_std_err"],
          "weights", r["level_weights"], "means", g["mean_diffs"], flush=True)
    assert np.allclose(r["level_weights"], [1.0, 44.0 / 45.0, 64.0 / 45.0], rtol=0, atol=1e-13)
    md = np.abs(g["mean_diffs"])
    assert 3.0 < md[1] / md[2] < 5.0, md  # weak order ~2
    bias_g = abs(g["price"] - fo)
    assert bias_g > 20.0 * g["price_std_err"], "the 4-step scheme's bias should be visible here"
    assert abs(r["price"] - fo) < 3.0 * r["price_std_err"] + 0.02 * bias_g


def test_ml2r_depth_and_accuracy(engine):
    """Adaptive depth: both estimators reach the Fourier price to within the
    RMSE target; ML2R stops at no more levels than Giles' MLMC."""
    m = make_model(1, [H1])
    fo = float(engine.fourier(m, EURO, capi.GREEK_NONE, S0, R, Q)["price"][0])
    eps = 3e-4 * fo
    out = {}
    for est in (capi.MLMC_GILES, capi.MLMC_ML2R):
        out[est] = engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_NONE,
                               mlmc_config(epsilon=eps, L_min=1, L_max=8, N0=20000, base_steps=1, seed=11,
                                           alpha=2.0, adaptive_rates=False, estimator=est,
                                           num_streams=1 << 16), S0, R, Q)
    g, r = out[capi.MLMC_GILES], out[capi.MLMC_ML2R]
    print("depth: giles", g["n_levels"], g["price"], "ml2r", r["n_levels"], r["price"], "fourier", fo, flush=True)
    assert r["n_levels"] <= g["n_levels"]
    for x in (g, r):
        assert abs(x["price"] - fo) < 3.0 * eps, (x["price"], fo, eps)


def test_richardson_romberg_is_two_level_ml2r(engine):
    m = make_model(1, [H1])
    r = engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_NONE,
                    mlmc_config(epsilon=0.01, L_min=1, L_max=1, N0=20000, base_steps=2, seed=3, alpha=2.0,
                                estimator=capi.MLMC_ML2R, num_streams=1 << 12), S0, R, Q)
    assert r["n_levels"] == 2
    assert np.allclose(r["level_weights"], [1.0, 4.0 / 3.0], rtol=0, atol=1e-14)
    assert abs(r["price"] - (r["mean_diffs"][0] + 4.0 / 3.0 * r["mean_diffs"][1])) < 1e-12 * r["price"]


def test_bad_estimator_refused(engine):
    m = make_model(1, [H1])
    from myquantlib_amd.pricing import EngineError
    with pytest.raises(EngineError):
        engine.mlmc(m, [capi.SCHEME_NV], EURO, capi.GREEK_NONE,
                    mlmc_config(epsilon=0.01, L_min=1, L_max=1, N0=2000, base_steps=2, estimator=7), S0, R, Q)
