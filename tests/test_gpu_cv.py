"""Learned martingale control variate on the GPU (no reference
implementation: checked against the FP64 numpy network in oracle/cv_mlp.py,
the engine's own plain estimator on the same paths, and Fourier prices).

  * the BF16 MFMA input gradient of the pricing kernel against FP64;
  * the FP32 training gradient against FP64 backpropagation;
  * the plain estimate of a CV run equals myql_mc_calculate on the same
    seed (same streams, partition and draws);
  * a trained CV lowers the variance, and its price agrees with the plain
    price and with Fourier: the control variate has zero mean.
"""
from __future__ import annotations

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, cv_config, make_model, mc_config
from oracle import cv_mlp
from oracle import fourier

pytestmark = pytest.mark.gpu
HESTON = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
S0, R, Q = 100.0, 0.02, 0.0


def test_mfma_input_gradient_matches_fp64(engine):
    p = cv_mlp.random_params(11).astype(np.float32)
    f = np.random.default_rng(12).normal(size=(200, 6)).astype(np.float32)
    print("cv input gradient: 200 rows", flush=True)
    g = engine.cv_input_grad(p, f)
    ref = cv_mlp.input_grad(p.astype(np.float64), f.astype(np.float64))[:, 1:4]
    # BF16 operands (8-bit mantissa) through three layers
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    assert err < 0.03, err


def test_training_gradient_matches_fp64(engine):
    p = cv_mlp.random_params(13).astype(np.float32)
    rng = np.random.default_rng(14)
    f = rng.normal(size=(128, 6)).astype(np.float32)
    t = rng.normal(size=128).astype(np.float32)
    print("cv training gradient: 128 rows", flush=True)
    g, loss = engine.cv_param_grad(p, f, t)
    ref, ref_loss = cv_mlp.param_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64))
    assert np.max(np.abs(g - ref)) <= 1e-4 * np.max(np.abs(ref))
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss


def _jobs(paths, seed, maturities=(0.5, 1.0), strikes=np.linspace(80, 120, 16), precision=capi.FP32):
    m = make_model(1, [HESTON])
    return [(m, [capi.SCHEME_NV], InstrumentSpec(strikes=strikes, maturity=T), capi.GREEK_NONE,
             mc_config(paths, int(round(50 * T)), seed=seed, precision=precision), S0, R, Q)
            for T in maturities]


def test_plain_estimate_of_cv_run_equals_engine(engine):
    jobs = _jobs(1 << 14, 42)
    cfg = cv_config(train_paths=1 << 12, steps=2, batch=256, points=4)
    print("cv train (2 steps) + price", flush=True)
    net = engine.cv_train(jobs, cfg)
    res = engine.mc_batch_cv(jobs, cfg, net)
    plain = engine.mc_batch(jobs)
    for a, b in zip(res, plain):
        np.testing.assert_allclose(a["plain_price"], b["price"], rtol=1e-12)
        np.testing.assert_allclose(a["plain_std_err"], b["price_std_err"], rtol=1e-9)


@pytest.fixture(scope="module")
def trained(engine):
    cfg = cv_config(train_paths=1 << 16, steps=400, batch=4096, points=8, lr=2e-3)
    print("cv train: 2^16 paths x 2 maturities, 400 Adam steps", flush=True)
    net = engine.cv_train(_jobs(1 << 10, 1), cfg)
    return cfg, net


def test_trained_cv_reduces_variance_and_is_unbiased(engine, trained):
    cfg, net = trained
    assert np.isfinite(net.final_loss)
    jobs = _jobs(1 << 16, 42)
    print("cv price: 2^16 paths x 2 maturities", flush=True)
    res = engine.mc_batch_cv(jobs, cfg, net)
    am = fourier.AffineModel(1, [HESTON])
    for (m, sch, inst, *_), r in zip(jobs, res):
        # the network must learn something: variance well below plain
        assert np.median(r["var_ratio"]) > 3.0, r["var_ratio"]
        ref = np.array([fourier.price(am, S0, K, inst.maturity, R, Q) for K in inst.strikes])
        z_fourier = (r["price"] - ref) / r["price_std_err"]
        assert np.max(np.abs(z_fourier)) < 4.5, z_fourier
        # P - CV against P on the same paths: Var(CV) <= (sd(P) + sd(P - CV))^2
        z_plain = (r["price"] - r["plain_price"]) / (r["price_std_err"] + r["plain_std_err"])
        assert np.max(np.abs(z_plain)) < 4.0, z_plain
