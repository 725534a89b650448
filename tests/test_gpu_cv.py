"""Learned martingale control variate on the GPU (no reference
implementation: checked against the FP64 numpy network in oracle/cv_mlp.py,
the engine's own plain estimator on the same paths, and Fourier prices).

  * the FP16 MFMA forward-mode tangents of the pricing kernel against FP64;
  * the FP32 training gradient against FP64 backpropagation;
  * the plain estimate of a CV run equals myql_mc_calculate on the same
    seed (same streams, partition and draws), and so do its Essential-mode
    delta and gamma;
  * trained control variates lower the variance and their prices agree with
    the plain prices of the same paths and with Fourier: every term of the
    estimator has zero mean (Heston NV; Bates NCI with Merton jumps, whose
    terminal jump is taken in expectation; Bates-Kou NV with vanilla calls
    and auto-smoothed cash digitals).
"""
from __future__ import annotations

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, cv_config, kou, make_model, mc_config, merton
from oracle import cv_mlp
from oracle import fourier

pytestmark = pytest.mark.gpu
HESTON = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
S0, R, Q = 100.0, 0.02, 0.0


def test_mfma_tangents_match_fp64(engine):
    p = cv_mlp.random_params(11).astype(np.float32)
    f = np.random.default_rng(12).normal(size=(200, 5)).astype(np.float32)
    print("cv tangents: 200 rows", flush=True)
    g = engine.cv_input_grad(p, f)
    ref = cv_mlp.tangents(p.astype(np.float64), f.astype(np.float64))
    # FP16 operands (11-bit mantissa; the activations and tangents are
    # converted rounding toward zero) through three products
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    print(f"max error / max |tangent|: {err:.2e}", flush=True)
    assert err < 0.01, err


def test_training_gradient_matches_fp64(engine):
    p = cv_mlp.random_params(13).astype(np.float32)
    rng = np.random.default_rng(14)
    f = rng.normal(size=(128, 5)).astype(np.float32)
    t = rng.normal(size=(128, 64)).astype(np.float32)
    print("cv training gradient: 128 rows", flush=True)
    g, loss = engine.cv_param_grad(p, f, t)
    ref, ref_loss = cv_mlp.param_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64))
    assert np.max(np.abs(g - ref)) <= 1e-4 * np.max(np.abs(ref))
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss


def _jobs(m, scheme, paths, seed, maturities=(0.5, 1.0), strikes=np.linspace(80, 120, 16), payoff=capi.PAYOFF_VANILLA,
          mode=capi.GREEK_NONE, steps_per_year=50, precision=capi.FP32):
    return [(m, [scheme], InstrumentSpec(payoff=payoff, strikes=strikes, maturity=T), mode,
             mc_config(paths, int(round(steps_per_year * T)), seed=seed, precision=precision), S0, R, Q)
            for T in maturities]


def test_plain_estimate_of_cv_run_equals_engine(engine):
    m = make_model(1, [HESTON])
    jobs = _jobs(m, capi.SCHEME_NV, 1 << 14, 42)
    cfg = cv_config(train_paths=1 << 12, steps=2, batch=256, points=4)
    print("cv train (2 steps) + price", flush=True)
    net = engine.cv_train(_jobs(m, capi.SCHEME_NV, 1024, 1), cfg)
    res = engine.mc_batch_cv(jobs, cfg, net)
    plain = engine.mc_batch(jobs)
    for a, b in zip(res, plain):
        np.testing.assert_allclose(a["plain_price"], b["price"], rtol=1e-12)
        np.testing.assert_allclose(a["plain_std_err"], b["price_std_err"], rtol=1e-9)


def test_essential_mode_plain_greeks_equal_engine(engine):
    """Auto-smoothed cash digitals in Essential mode: the CV run's plain
    price, delta and gamma equal myql_mc_calculate's on the same seed."""
    jk, jp = kou(1.0, 0.4, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jk, jump=jp)
    jobs = _jobs(m, capi.SCHEME_NV, 1 << 14, 42, payoff=capi.PAYOFF_CASH, mode=capi.GREEK_ESSENTIAL)
    cfg = cv_config(train_paths=1 << 12, steps=2, batch=256, points=4)
    print("cv essential digitals: train (2 steps) + price", flush=True)
    net = engine.cv_train(jobs, cfg)
    res = engine.mc_batch_cv(jobs, cfg, net)
    plain = engine.mc_batch(jobs)
    for a, b in zip(res, plain):
        np.testing.assert_allclose(a["plain_price"], b["price"], rtol=1e-12)
        for f in ("delta", "gamma", "delta_std_err", "gamma_std_err"):
            np.testing.assert_allclose(a[f], b[f], rtol=1e-9, atol=1e-14)


def _check_unbiased(res, ref=None, z_max=4.5):
    for r in res:
        # CV against the plain estimate of the same paths:
        # Var(P - CV) <= (sd(P) + sd(Y))^2
        z_plain = (r["price"] - r["plain_price"]) / (r["price_std_err"] + r["plain_std_err"])
        assert np.max(np.abs(z_plain)) < 4.0, z_plain
    if ref is not None:
        for r, f in zip(res, ref):
            z = (r["price"] - f) / r["price_std_err"]
            assert np.max(np.abs(z)) < z_max, z


def _fourier(engine, m, jobs):
    return [engine.fourier(m, inst, capi.GREEK_NONE, S0, R, Q)["price"] for (_m, _s, inst, *_r) in jobs]


def test_trained_cv_heston(engine):
    m = make_model(1, [HESTON])
    cfg = cv_config(train_paths=1 << 16, steps=400, batch=4096, points=8, lr=2e-3)
    print("cv heston: train 2^16 paths x 2 maturities, 400 Adam steps", flush=True)
    net = engine.cv_train(_jobs(m, capi.SCHEME_NV, 1 << 10, 1), cfg)
    assert np.isfinite(net.final_loss)
    jobs = _jobs(m, capi.SCHEME_NV, 1 << 16, 42)
    res = engine.mc_batch_cv(jobs, cfg, net)
    for r in res:
        assert np.median(r["var_ratio"]) > 5.0, r["var_ratio"]
    _check_unbiased(res, _fourier(engine, m, jobs))


def test_trained_cv_bates_nci_merton(engine):
    """The headline strip's model at reduced size: SchemeNCI (zero-mean
    variance increments from the table rows' means) and Merton jumps (the
    terminal jump in conditional expectation)."""
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)
    strikes = np.linspace(70, 130, 16)
    cfg = cv_config(train_paths=1 << 16, steps=400, batch=4096, points=8, lr=2e-3)
    print("cv bates nci: train 2^16 paths x 2 maturities, 400 Adam steps", flush=True)
    net = engine.cv_train(_jobs(m, capi.SCHEME_NCI, 1 << 10, 7, strikes=strikes, steps_per_year=100), cfg)
    jobs = _jobs(m, capi.SCHEME_NCI, 1 << 16, 42, strikes=strikes, steps_per_year=100)
    res = engine.mc_batch_cv(jobs, cfg, net)
    for r in res:
        assert np.median(r["var_ratio"]) > 4.0, r["var_ratio"]
    _check_unbiased(res, _fourier(engine, m, jobs))


def test_trained_cv_bates_kou_vanilla_and_smoothed_digitals(engine):
    """Configuration 5's model (Bates-Kou, SchemeNV) at reduced size: vanilla
    calls (against Fourier) and auto-smoothed cash digitals in Essential
    mode (against the plain estimate of the same paths, which equals
    myql_mc_calculate)."""
    jk, jp = kou(1.0, 0.4, 10.0, 5.0)
    m = make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jk, jump=jp)
    strikes = np.linspace(70, 130, 16)
    cfg = cv_config(train_paths=1 << 16, steps=400, batch=4096, points=8, lr=2e-3)
    for payoff, mode in ((capi.PAYOFF_VANILLA, capi.GREEK_NONE), (capi.PAYOFF_CASH, capi.GREEK_ESSENTIAL)):
        print(f"cv bates-kou payoff {payoff}: train + price", flush=True)
        net = engine.cv_train(_jobs(m, capi.SCHEME_NV, 1 << 10, 7, strikes=strikes, payoff=payoff, mode=mode,
                                    steps_per_year=100), cfg)
        jobs = _jobs(m, capi.SCHEME_NV, 1 << 16, 42, strikes=strikes, payoff=payoff, mode=mode, steps_per_year=100)
        res = engine.mc_batch_cv(jobs, cfg, net)
        plain = engine.mc_batch(jobs)
        for a, b in zip(res, plain):
            np.testing.assert_allclose(a["plain_price"], b["price"], rtol=1e-12)
            assert np.median(a["var_ratio"]) > 1.5, a["var_ratio"]
        _check_unbiased(res, _fourier(engine, m, jobs) if payoff == capi.PAYOFF_VANILLA else None)
