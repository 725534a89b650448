"""Launch shapes and the autotuner (MYQL_FLAG_AUTOTUNE) on the GPU.

A launch shape -- streams per work unit (block size), unroll of the strike
loop, persistent grid factor -- changes how the same streams, draws and
paths are scheduled, and for the block size the FP64 summation order of the
per-unit partial sums; it never changes a path. So every compiled shape and
the autotuner's choice must reproduce the default launch's prices and
Greeks to FP64 summation noise (1e-12 relative to each field's scale), and
a tuned context must repeat itself bit for bit.
"""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Engine, Heston, InstrumentSpec, make_model, mc_config, merton

pytestmark = pytest.mark.gpu

FIELDS = ["price", "price_std_err", "delta", "gamma", "vega", "theta", "rho", "vanna", "charm"]


def strip_jobs(prec, flags=0):
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)
    out = []
    for T in (0.25, 1.0):
        inst = InstrumentSpec(strikes=np.linspace(70, 130, 24), maturity=T)
        # 3000 streams: the last unit of each job is partly empty at both block sizes
        cfg = mc_config(1 << 16, int(100 * T), seed=42, precision=prec, num_streams=3000, flags=flags)
        out.append((m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL, cfg, 100.0, 0.02, 0.0))
    return out


def close(a, b, tol=1e-12):
    for ra, rb in zip(a, b):
        for f in FIELDS:
            x, y = np.atleast_2d(np.asarray(ra[f], float)), np.atleast_2d(np.asarray(rb[f], float))
            scale = max(float(np.max(np.abs(y))), 1e-300)
            assert np.max(np.abs(x - y)) <= tol * scale, f


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_every_launch_shape_matches_default(engine, prec):
    ref = engine.mc_batch(strip_jobs(prec))
    tol = 1e-12 if prec == capi.FP64 else 1e-9
    eng = Engine(0)
    try:
        for blk, unr, gf, w4 in [(256, 1, 4, 0), (128, 4, 4, 0), (128, 1, 1, 0), (256, 4, 1, 0), (64, 4, 4, 0),
                                 (64, 1, 2, 0), (64, 4, 4, 1), (64, 1, 1, 1)]:
            eng.set_launch_shape(prec, 1, 5, False, blk, unr, gf, bool(w4))
            close(eng.mc_batch(strip_jobs(prec)), ref, tol)
        with pytest.raises(Exception):
            eng.set_launch_shape(prec, 1, 5, False, 32, 4, 4)   # not compiled
        with pytest.raises(Exception):
            eng.set_launch_shape(prec, 1, 5, False, 128, 4, 4, True)   # four-wave variant: 64 only
    finally:
        eng.close()


def test_autotuned_strip_matches_default_and_repeats(engine):
    ref = engine.mc_batch(strip_jobs(capi.FP64))
    eng = Engine(0)
    try:
        assert eng.tuning_report() == ""
        a = eng.mc_batch(strip_jobs(capi.FP64, capi.FLAG_AUTOTUNE))
        rep = eng.tuning_report()
        assert rep.startswith("fp64 nf=1 nv=5 european: block"), rep
        assert rep.count(" ms") == 16, rep   # 8 shapes x 2 grid factors
        b = eng.mc_batch(strip_jobs(capi.FP64, capi.FLAG_AUTOTUNE))
        assert eng.tuning_report() == rep   # tuned once per context
        close(a, ref)
        for ra, rb in zip(a, b):
            for f in FIELDS:
                assert np.array_equal(np.asarray(ra[f]), np.asarray(rb[f])), f
    finally:
        eng.close()


def _case(kind):
    h1, h2 = Heston(2.0, 0.04, 0.3, -0.7, 0.04), Heston(1.0, 0.03, 0.9, -0.5, 0.02)
    if kind == "asian_lockstep":      # 1 factor, path-dependent lockstep (7 slots)
        m, sch = make_model(1, [h1]), [capi.SCHEME_NV]
        inst = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[90.0, 100.0, 110.0], maturity=1.0)
        mode, key = capi.GREEK_FULL, (1, 7, True)
    elif kind == "barrier_replay":    # 2 factors, Exact: one slot per replay
        m, sch = make_model(2, [h1, h2]), [capi.SCHEME_NV, capi.SCHEME_EXACT]
        inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=80.0, maturity=1.0,
                              barrier_dir=capi.BARRIER_DOWN, barrier_action=capi.BARRIER_OUT)
        mode, key = capi.GREEK_NONE, (2, 1, True)
    else:                             # 0 factors, European replay
        m, sch = make_model(0, [], vol=0.2), []
        inst = InstrumentSpec(strikes=[80.0, 100.0, 120.0], maturity=1.0)
        mode, key = capi.GREEK_NONE, (0, 1, False)
    cfg = mc_config(1 << 15, 60, seed=5, num_streams=1500)
    return (m, sch, inst, mode, cfg, 100.0, 0.02, 0.0), key


@pytest.mark.parametrize("kind", ["asian_lockstep", "barrier_replay", "european_0f"])
def test_other_kernels_at_every_block_size(engine, kind):
    job, (nf, nv, pd) = _case(kind)
    ref = engine.mc_batch([job])
    fields = FIELDS if job[3] == capi.GREEK_FULL else ["price", "price_std_err"]
    eng = Engine(0)
    try:
        for blk in (128, 64):
            eng.set_launch_shape(capi.FP64, nf, nv, pd, blk, 4, 4)
            got = eng.mc_batch([job])
            for f in fields:
                x, y = np.atleast_2d(np.asarray(got[0][f], float)), np.atleast_2d(np.asarray(ref[0][f], float))
                assert np.max(np.abs(x - y)) <= 1e-12 * max(float(np.max(np.abs(y))), 1e-300), (blk, f)
        with pytest.raises(Exception):
            eng.set_launch_shape(capi.FP64, nf, nv, pd, 64, 1, 4)   # unroll 1 only for lockstep Europeans
    finally:
        eng.close()
