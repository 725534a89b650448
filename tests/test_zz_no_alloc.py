"""myql_mc_calculate_noalloc (no counterpart in the reference: the entry
point for callers that cannot allocate device memory) against
myql_mc_calculate through the C-ABI: the same kernels on the same streams,
so every result field is identical, on two calls in a row, the second with
a smaller input (fewer paths, streams and strikes), which also leaves the
shared scratch holding the first call's contents. Sizes beyond the scratch
and SchemeNCI are refused before any device work."""
import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import EngineError, Heston, InstrumentSpec, make_model, mc_config, merton

pytestmark = pytest.mark.gpu

FIELDS = ("price", "price_std_err", "delta", "delta_std_err", "gamma", "gamma_std_err", "vega", "theta", "rho",
          "vanna", "charm", "charm_std_err")


def _same(a, b):
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert np.array_equal(x, y, equal_nan=True), (f, x, y)


def test_noalloc_equals_mc_calculate_on_two_calls(engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [Heston(2.0, 0.04, 0.3, -0.7, 0.04)], jump_kind=jk, jump=jp)
    calls = [
        (InstrumentSpec(strikes=np.linspace(70, 130, 40), maturity=1.0), capi.GREEK_FULL,
         mc_config(1 << 16, 50, seed=42, precision=capi.FP32, num_streams=8192)),
        (InstrumentSpec(strikes=[90.0, 100.0, 110.0], maturity=0.5), capi.GREEK_ESSENTIAL,
         mc_config(1 << 12, 20, seed=7, precision=capi.FP64, num_streams=1024)),
    ]
    for i, (inst, mode, cfg) in enumerate(calls):
        print(f"noalloc call {i}: {len(inst.strikes)} strikes, {cfg.num_paths} paths", flush=True)
        want = engine.mc(m, [capi.SCHEME_NV], inst, mode, cfg, 100.0, 0.02, 0.0)
        got = engine.mc_noalloc(m, [capi.SCHEME_NV], inst, mode, cfg, 100.0, 0.02, 0.0)
        _same(got, want)


def test_noalloc_path_dependent_and_two_factor(engine):
    m = make_model(2, [Heston(2.0, 0.04, 0.3, -0.7, 0.04), Heston(1.0, 0.03, 0.25, -0.5, 0.02)])
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[95.0, 105.0], barrier=130.0, maturity=1.0)
    cfg = mc_config(1 << 14, 64, seed=3, precision=capi.FP32, num_streams=4096)
    want = engine.mc(m, [capi.SCHEME_NV, capi.SCHEME_EXACT], inst, capi.GREEK_FULL, cfg, 100.0, 0.02, 0.0)
    got = engine.mc_noalloc(m, [capi.SCHEME_NV, capi.SCHEME_EXACT], inst, capi.GREEK_FULL, cfg, 100.0, 0.02, 0.0)
    _same(got, want)


def test_noalloc_refuses_what_does_not_fit(engine):
    m = make_model(1, [Heston(2.0, 0.04, 0.3, -0.7, 0.04)])
    inst = InstrumentSpec(strikes=np.linspace(70, 130, 64), maturity=1.0)
    big = mc_config(1 << 22, 10, seed=1, num_streams=1 << 20)  # 32 MiB of states + 160 MiB of accumulators
    with pytest.raises(EngineError) as e:
        engine.mc_noalloc(m, [capi.SCHEME_NV], inst, capi.GREEK_FULL, big, 100.0, 0.02, 0.0)
    assert e.value.code == capi.E_INVALID_ARG and "scratch" in str(e.value)
    nci = make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)])
    with pytest.raises(EngineError) as e:
        engine.mc_noalloc(nci, [capi.SCHEME_NCI], inst, capi.GREEK_NONE, mc_config(1 << 10, 10), 100.0, 0.02, 0.0)
    assert e.value.code == capi.E_UNSUPPORTED
    # the context still prices normally afterwards, with and without the scratch
    small = mc_config(1 << 12, 10, seed=2, num_streams=512)
    _same(engine.mc_noalloc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, small, 100.0, 0.02, 0.0),
          engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, small, 100.0, 0.02, 0.0))
