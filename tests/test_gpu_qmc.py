"""Randomized quasi-Monte Carlo (MYQL_STREAMS_SOBOL) on the GPU.

No reference implementation (README.md:332-335 lists QMC as future work).
Checked:
 - the coordinates the path kernels draw equal oracle/sobol.py's digitally
   shifted Sobol' points bit for bit (the restatement itself is pinned by
   tests/test_sobol.py), so the pricing paths, which are the pseudo-random
   paths' code on these uniforms, are what they should be;
 - unbiasedness: prices within 3.5-4 replicate standard errors of the
   Gil-Pelaez prices (Black-Scholes, Bates NCI with Full Greeks) or of
   pseudo-random MC at the same steps (Heston NV, a barrier under Bates);
 - the point of it: at equal paths the replicate standard error is far
   below the pseudo-random one where the integrand is smooth and
   low-dimensional.
"""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import EngineError, Heston, InstrumentSpec, make_model, mc_config, merton
from oracle import sobol

pytestmark = pytest.mark.gpu

S0, R, Q = 100.0, 0.02, 0.0


def test_device_coordinates_equal_the_restatement(engine):
    dims, n_points, n_streams = 40, 4096, 1024  # 4 wave steps per wave
    for seed, rep in ((42, 0), (42, 3), (7, 1)):
        got = engine.qmc_points(seed, rep, dims, n_points, n_streams)
        V = sobol.direction_numbers(dims)
        shift = sobol.replicate_shift(seed, rep, dims)
        n_it = n_points // n_streams
        for i in (0, 1, 63, 64, 65, 1000, 2047, 4095):
            h, lane = divmod(i, 64)
            for d in (0, 1, 2, 17, 39):
                assert int(got[i, d]) == sobol.engine_point(V, shift, h, lane, d), (seed, rep, i, d)
        # whole columns as sets: each wave walks its own n_it steps
        assert n_it == 4
        want = np.array([[sobol.engine_point(V, shift, h, lane, d) for d in range(3)]
                         for h in range(n_points // 64) for lane in range(64)], dtype=np.uint32)
        assert np.array_equal(got[:, :3], want)


def _z(res, ref, f="price"):
    x, e = np.asarray(res[f], float), np.asarray(res[f + "_std_err"], float)
    return (x - np.asarray(ref[f], float)) / e


def test_black_scholes_call_unbiased_and_sharper(engine):
    m = make_model(0, vol=0.25)
    inst = InstrumentSpec(strikes=[80.0, 100.0, 120.0], maturity=1.0)
    fo = engine.fourier(m, inst, capi.GREEK_NONE, S0, R, Q)
    q = engine.mc(m, [], inst, capi.GREEK_NONE,
                  mc_config(1 << 20, 1, seed=42, stream_layout=capi.STREAMS_SOBOL, num_streams=1 << 14), S0, R, Q)
    p = engine.mc(m, [], inst, capi.GREEK_NONE, mc_config(1 << 20, 1, seed=42, num_streams=1 << 14), S0, R, Q)
    z = _z(q, fo)
    print("bs qmc", q["price"], q["price_std_err"], "plain se", p["price_std_err"], "z", z, flush=True)
    assert np.all(np.abs(z) < 3.5)
    # one coordinate: the randomized net's error is orders below Monte Carlo's
    assert np.all(np.asarray(q["price_std_err"]) < 0.05 * np.asarray(p["price_std_err"]))


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_heston_strip_unbiased_and_sharper(engine, prec):
    """Heston NV, 8 steps (16 coordinates): against pseudo-random MC at the
    same steps (independent streams, 4x the paths), so the scheme's
    discretisation bias is common to both."""
    m = make_model(1, [Heston(2.0, 0.04, 0.3, -0.7, 0.04)])
    inst = InstrumentSpec(strikes=np.linspace(80.0, 120.0, 9), maturity=1.0)
    cfg_q = mc_config(1 << 20, 8, seed=42, precision=prec, stream_layout=capi.STREAMS_SOBOL, num_streams=1 << 14)
    q = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL, cfg_q, S0, R, Q)
    p1 = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                   mc_config(1 << 20, 8, seed=42, precision=prec, num_streams=1 << 14), S0, R, Q)
    p4 = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                   mc_config(1 << 22, 8, seed=43, precision=prec, num_streams=1 << 16), S0, R, Q)
    ratio = np.asarray(p1["price_std_err"]) / np.asarray(q["price_std_err"])
    for f in ("price", "delta"):
        z = (np.asarray(q[f]) - np.asarray(p4[f])) / np.hypot(q[f + "_std_err"], p4[f + "_std_err"])
        print("heston qmc", f, "z", z, flush=True)
        assert np.all(np.abs(z) < 3.5), (f, z)
    print("heston qmc se ratio", ratio, flush=True)
    assert np.median(ratio) > 2.0
    assert q["total_steps"] == p1["total_steps"]


def test_bates_nci_full_greeks(engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(strikes=np.linspace(80.0, 120.0, 8), maturity=0.5)
    fo = engine.fourier(m, inst, capi.GREEK_FULL, S0, R, Q)
    q = engine.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL,
                  mc_config(1 << 19, 50, seed=9, precision=capi.FP32, stream_layout=capi.STREAMS_SOBOL,
                            num_streams=1 << 13, qmc_replicates=8), S0, R, Q)
    for f in ("price", "delta", "vega"):
        z = _z(q, fo, f) if f != "vega" else (np.asarray(q["vega"][0]) - np.asarray(fo["vega"][0])) / np.asarray(
            q["vega_std_err"][0])
        print("bates qmc", f, z, flush=True)
        assert np.all(np.isfinite(z)) and np.all(np.abs(z) < 4.0), (f, z)


def test_path_dependent_with_merton_jumps(engine):
    """A barrier under Bates (per-step Merton jumps: count and size
    coordinates every step) agrees with pseudo-random MC."""
    jk, jp = merton(0.3, -0.05, 0.1)
    m = make_model(1, [Heston(2.0, 0.04, 0.3, -0.7, 0.04)], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=120.0, maturity=0.5,
                          barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)
    q = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                  mc_config(1 << 19, 25, seed=3, stream_layout=capi.STREAMS_SOBOL, num_streams=1 << 13), S0, R, Q)
    p = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL,
                  mc_config(1 << 21, 25, seed=4, num_streams=1 << 15), S0, R, Q)
    for f in ("price", "delta"):
        se = np.hypot(q[f + "_std_err"][0], p[f + "_std_err"][0])
        z = (q[f][0] - p[f][0]) / se
        print("barrier qmc", f, q[f][0], p[f][0], z, flush=True)
        assert abs(z) < 3.5


def test_refusals(engine):
    m = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)])
    inst = InstrumentSpec(strikes=[100.0], maturity=1.0)
    cfg = mc_config(1 << 12, 10, stream_layout=capi.STREAMS_SOBOL)
    with pytest.raises(EngineError) as e:
        engine.mc(m, [capi.SCHEME_EXACT], inst, capi.GREEK_NONE, cfg, S0, R, Q)
    assert e.value.code == capi.E_UNSUPPORTED
    with pytest.raises(EngineError) as e:
        engine.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_NONE,
                  mc_config(1 << 12, 400, stream_layout=capi.STREAMS_SOBOL), S0, R, Q)  # 1,200 coordinates
    assert e.value.code == capi.E_UNSUPPORTED
    with pytest.raises(EngineError) as e:
        engine.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_NONE,
                  mc_config(1 << 12, 10, stream_layout=capi.STREAMS_SOBOL, qmc_replicates=1), S0, R, Q)
    assert e.value.code == capi.E_INVALID_ARG
