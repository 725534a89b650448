"""GPU parity: the HIP engine against the CPU FP64 oracle on identical draws.

Every call goes through the C-ABI (libmyql_hip.so). Tolerances:
  integer stream and Poisson counts: bit-exact;
  FP64 paths: 1e-9 relative on prices and Greeks (the engine and the oracle
    differ only by FMA contraction and libm-vs-device ulps);
  FP32 paths (sums and central differences in FP64): stated per test.
"""
import numpy as np
import pytest
from scipy import stats

from myquantlib_amd import capi
from myquantlib_amd.pricing import (Heston, InstrumentSpec, make_model, mc_config, merton, kou,
                                    mlmc_config)

pytestmark = pytest.mark.gpu

H1 = Heston(2.0, 0.04, 0.3, -0.7, 0.04)          # Feller-compliant (NV allowed)
H_HI = Heston(2.0, 0.04, 1.0, -0.7, 0.04)        # Feller violated (NCI / Exact)
H2 = Heston(1.0, 0.03, 0.25, -0.5, 0.02)


def rel_err(a, b, floor=1e-12):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


# Second-order and cross Greeks are central differences amplified by 1/h^2
# (gamma), 1/(2 dv 2h) (vanna) or 1/(2 dT 2h) (charm): FMA-level rounding
# differences of the payoffs show up there at 1e-9..1e-7 of the field's scale,
# so they are held to the 1e-6 bar; prices and first-order Greeks to 1e-9.
SECOND_ORDER = {"gamma", "vanna", "charm", "gamma_std_err", "vanna_std_err", "charm_std_err"}


def compare(gpu, cpu, fields, tol, floor=None):
    for f in fields:
        g, c = np.asarray(gpu[f], dtype=float), np.asarray(cpu[f], dtype=float)
        mask = np.isfinite(c)
        assert np.all(np.isfinite(g[mask])), f
        # errors are measured against the field's scale across the strip, so
        # entries that are zero up to rounding (deep ITM/OTM gammas) count
        # by their absolute size
        fl = floor if floor is not None else max(1e-12, 1e-3 * float(np.max(np.abs(c[mask]), initial=0.0)))
        t = tol
        if f in SECOND_ORDER:
            # exactly-zero second differences (linear payoffs, homogeneous
            # lookbacks) carry ~1e-11 of rounding residue: judged against 1e-4
            t = max(tol, 1e-6)
            fl = max(fl, 1e-4)
        e = rel_err(g[mask], c[mask], fl)
        assert e <= t, f"{f}: rel err {e:.3e} > {t:.1e}\n gpu={g}\n cpu={c}"


ESS = ["price", "price_std_err", "delta", "gamma"]
FULL = ESS + ["vega", "theta", "rho", "vanna", "charm"]


@pytest.mark.parametrize("layout", [capi.STREAMS_JUMP, capi.STREAMS_SEEDED])
def test_stream_draws_bit_exact(engine, oracle, layout):
    n_streams, n_per = 1000, 40
    gpu = engine.stream_draws(42, layout, n_streams, n_per)
    for s in [0, 1, 2, 3, 7, 255, 256, 511, 999]:
        st = oracle.stream_state(42, layout, s)
        cpu = oracle.next(st, n_per)
        assert np.array_equal(gpu[s], cpu), f"stream {s}"


def test_normcdfinv_device(engine, oracle):
    p = np.concatenate([np.linspace(1e-12, 1 - 1e-12, 20001), 10.0 ** -np.arange(1, 18),
                        1 - 10.0 ** -np.arange(1, 16), [0.02425, 0.97575]])
    o64, o32 = engine.normcdfinv(p)
    ref = np.array([oracle.normcdfinv(x) for x in p])
    err = np.abs(o64 - ref) / np.maximum(np.abs(ref), 1e-6)
    i = int(np.argmax(err))
    # device log/sqrt and FMA contraction differ from glibc by ulps
    assert err[i] < 1e-12, (p[i], o64[i], ref[i], err[i])
    assert np.max(np.abs(o32 - ref)) < 2e-5 * np.maximum(1.0, np.abs(ref)).max()


def test_poissinv_device_exact(engine, oracle):
    rng = np.random.default_rng(1)
    lam = np.concatenate([rng.uniform(0, 4, 3000), rng.uniform(4, 40, 3000),
                          10 ** rng.uniform(1, 4, 2000), [0.0, 3.999, 4.0, 4.001]])
    u = rng.uniform(0, 1, lam.size)
    u[:50] = 1e-15
    u[50:100] = 1 - 1e-12
    out = engine.poissinv(lam, u)
    ref = np.array([oracle.poissinvcdf(a, b) for a, b in zip(lam, u)])
    assert np.array_equal(out, ref)


def _near_cdf_uniforms(lam, rng, n_per):
    """Uniforms at distances 1e-7 .. 1e-3 either side of F(n) for the counts
    around the mean: where the FP32 first pass must defer to the FP64 method."""
    ks = np.arange(max(0, int(lam - 3 * np.sqrt(lam) - 2)), int(lam + 3 * np.sqrt(lam) + 3))
    c = stats.poisson.cdf(ks, lam)
    d = np.concatenate([10.0 ** rng.uniform(-7, -3, n_per), -(10.0 ** rng.uniform(-7, -3, n_per))])
    u = (c[:, None] + d[None, :]).ravel()
    return u[(u > 0) & (u < 1)]


def test_poisson_first_pass_counts_equal_giles(engine, oracle):
    """The path kernels' Poisson counts (FP32 first pass, FP64 method for the
    undecided) against the oracle's Giles inversion, on random (lambda, u)
    and on uniforms placed within 1e-7..1e-3 of the CDF values, for one slot
    and for five slots sharing u (base and bumps of 0, +-1e-5 and +-1e-2)."""
    rng = np.random.default_rng(11)
    lam = np.concatenate([rng.uniform(0, 4, 20000), rng.uniform(4, 40, 20000), 10 ** rng.uniform(1.6, 3, 5000),
                          [4.0, 4.0000001, 10.0, 20.0]])
    u = rng.uniform(0, 1, lam.size)
    u[:30] = 0.0
    u[30:60] = 1e-300
    lams, us = [lam], [u]
    for l0 in (0.3, 2.0, 3.99, 4.01, 6.5, 9.0, 12.0, 25.0, 60.0, 250.0):
        uu = _near_cdf_uniforms(l0, rng, 40)
        lams.append(np.full(uu.size, l0) * np.exp(rng.uniform(-0.05, 0.05, uu.size)))
        us.append(uu)
    # Giles' tail branch (|W| >= 3, run in FP32 by the wave-packed pass):
    # uniforms down to 1e-19 and up to 1 - 1e-15, and placed near the CDF
    # values 3-8 standard deviations out
    for l0 in (4.5, 9.0, 9.8, 15.0, 40.0, 150.0, 900.0, 2500.0):
        ll = np.full(400, l0) * np.exp(rng.uniform(-0.05, 0.05, 400))
        uu = np.concatenate([10 ** rng.uniform(-19, -2.9, 200), 1.0 - 10 ** rng.uniform(-15, -2.9, 200)])
        lams.append(ll)
        us.append(uu)
        sd = np.sqrt(l0)
        hi = np.arange(int(l0 + 3 * sd), int(l0 + 8 * sd) + 2)
        lo = np.arange(max(0, int(l0 - 8 * sd)), max(1, int(l0 - 3 * sd)))
        d = 10.0 ** rng.uniform(-7, -4, (hi.size + lo.size, 6)) * np.where(
            rng.uniform(size=(hi.size + lo.size, 6)) < 0.5, -1, 1)
        # relative offsets from the tail probability on either side
        uu = np.concatenate([(1.0 - stats.poisson.sf(hi, l0)[:, None] * (1.0 + d[:hi.size])).ravel(),
                             (stats.poisson.cdf(lo, l0)[:, None] * (1.0 + d[hi.size:])).ravel()])
        uu = uu[(uu > 0) & (uu < 1)]
        lams.append(np.full(uu.size, l0))
        us.append(uu)
    lam, u = np.concatenate(lams), np.concatenate(us)
    ref = np.array([oracle.poissinvcdf(a, b) for a, b in zip(lam, u)]).astype(np.int32)
    got = engine.poisson_counts(lam, u)[:, 0]
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, (lam[bad[:5]], u[bad[:5]], got[bad[:5]], ref[bad[:5]])
    bumps = np.array([0.0, 1e-5, -1e-5, 1e-2, -1e-2])
    L5 = lam[:, None] * (1.0 + bumps[None, :])
    got5 = engine.poisson_counts(L5, u)
    ref5 = np.array([[oracle.poissinvcdf(a, b) for a in row] for row, b in zip(L5, u)]).astype(np.int32)
    bad = np.argwhere(got5 != ref5)
    assert bad.size == 0, [(L5[i, k], u[i], got5[i, k], ref5[i, k]) for i, k in bad[:5]]


def test_poissinv_device_integer_intensities(engine, oracle):
    """Integer intensities put the count X = lam in Giles' Temme correction
    with X / lam = 1 exactly (jump means lambda T = 10, 20, ...). u around
    F(lam - 1) is where a contracted 1 - X/lam went negative on the device
    and returned lam instead of lam - 1."""
    rng = np.random.default_rng(2)
    for lam in (10.0, 11.0, 20.0, 37.0, 100.0):
        c = stats.poisson.cdf(lam - 1, lam)
        u = np.concatenate([c + rng.uniform(-2e-3, 2e-3, 3000), rng.uniform(0, 1, 1000)])
        out = engine.poissinv(np.full(u.size, lam), u)
        ref = np.array([oracle.poissinvcdf(lam, b) for b in u])
        assert np.array_equal(out, ref), lam
        ppf = stats.poisson.ppf(u, lam)
        ok = np.abs(stats.poisson.cdf(ppf, lam) - u) > 1e-12
        assert np.array_equal(out[ok], ppf[ok]), lam


def test_chi2_quantile_device(engine, oracle):
    rng = np.random.default_rng(2)
    # up to the last NCI table row's d + 2N (N = 1023)
    dof = np.concatenate([rng.uniform(0.05, 2, 500), rng.uniform(2, 200, 500), rng.uniform(200, 2100, 1000)])
    p = rng.uniform(0, 1, dof.size)
    p[:20] = 0.999
    p[20:40] = 1 - 1e-9
    p[40:60] = 1e-6
    # the first NCI grid interval (inverted exactly for rows N >= 2)
    p[500:700] = np.exp(rng.uniform(np.log(2.0 ** -64), np.log(1e-4), 200))
    p[1000:1200] = np.exp(rng.uniform(np.log(2.0 ** -64), np.log(1e-4), 200))
    out = engine.chi2_quantile(dof, p)
    ref = np.array([oracle.chi2_quantile(a, b) for a, b in zip(dof, p)])
    assert rel_err(out, ref, 1e-300) < 1e-12


def _run_both(engine, oracle, model, schemes, inst, mode, cfg, S0=100.0, r=0.02, q=0.0):
    g = engine.mc(model, schemes, inst, mode, cfg, S0, r, q)
    c = oracle.mc(model, schemes, inst, mode, cfg, S0, r, q)
    return g, c


def test_heston_nv_essential_fp64(engine, oracle):
    """Configuration 1 shape (Heston, NV, European call, Essential) at 2^16 paths."""
    m = make_model(1, [H1])
    inst = InstrumentSpec(strikes=[100.0], maturity=1.0)
    cfg = mc_config(1 << 16, 50, seed=42, num_streams=4096)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NV], inst, capi.GREEK_ESSENTIAL, cfg)
    compare(g, c, ESS, 1e-9)
    assert g["total_steps"] == (1 << 16) * 50


@pytest.mark.parametrize("prec,tol,flags", [(capi.FP64, 1e-9, 0), (capi.FP32, 2e-3, 0),
                                            (capi.FP64, 1e-9, capi.FLAG_REFERENCE_NCI_ROWS)])
def test_bates_nci_full_strip(engine, oracle, prec, tol, flags):
    """Configuration 2 shape (Bates, NCI with Feller violated, Merton, Full
    CRN Greeks on a strike strip) at reduced size; with the reference's NCI
    row rule the exact chi2 inversion runs on ~4% of steps."""
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [H_HI], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(strikes=np.linspace(70, 130, 16), maturity=0.5)
    cfg = mc_config(1 << 14, 50, seed=42, precision=prec, num_streams=1024, flags=flags)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL, cfg)
    if prec == capi.FP64:
        compare(g, c, FULL, tol)
    else:
        # FP32 paths: prices and first-order Greeks against the FP64 oracle
        compare(g, c, ["price", "delta"], tol)


# FP32 paths against the FP64 oracle on identical draws, every Full field:
# max over options (price >= 0.1) of |fp32 - fp64| / max(|fp64|, its SE).
# The maturity-bumped variants behind theta and charm are stepped in FP64
# inside the FP32 kernel from the FP64 chi-squared rows, so their paths are
# the FP64 ones (theta and charm differ by ~1e-4 of SE at the full strip).
# The base and v0-bumped paths keep FP32 variance: where FP32 rounding of a
# table value moves a Poisson count across a CDF boundary their paths part
# ways from the FP64 ones (noise, largest in the most amplified estimators:
# gamma and vanna).
FP32_FULL_BOUNDS = {"price": 1e-4, "delta": 1e-4, "rho": 1e-4, "vega": 1e-3, "theta": 0.01, "gamma": 0.05,
                    "vanna": 0.05, "charm": 0.01}


def test_bates_nci_full_strip_fp32_every_field(engine, oracle):
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [H_HI], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(strikes=np.linspace(70, 130, 16), maturity=1.0)
    print("fp32 full strip vs fp64 oracle: 2^16 paths, 100 steps", flush=True)
    g = engine.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL,
                  mc_config(1 << 16, 100, seed=42, precision=capi.FP32, num_streams=1024), 100.0, 0.02, 0.0)
    c = oracle.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL,
                  mc_config(1 << 16, 100, seed=42, precision=capi.FP64, num_streams=1024), 100.0, 0.02, 0.0)
    ok = c["price"] >= 0.1
    errs = {}
    for f, bound in FP32_FULL_BOUNDS.items():
        x, y, se = (np.asarray(a, float).reshape(-1)[:16] for a in (g[f], c[f], c[f + "_std_err"]))
        errs[f] = float(np.max(np.abs(x - y)[ok] / np.maximum(np.abs(y), se)[ok]))
    bad = {f: e for f, e in errs.items() if not e <= FP32_FULL_BOUNDS[f]}
    assert not bad, (bad, errs)


@pytest.mark.parametrize("bump", [1e-3, 3e-2])
def test_bates_nci_bumped_slot_counts(engine, oracle, bump):
    """The lockstep kernel gives each bumped Greek slot (v0 +-, T +-) its
    Poisson count from slot 0's inversion -- CDF margins moved by the
    intensity change to second order, slot 0's asymptotic estimate, or a
    rerun of the asymptotic branch -- and only falls back to the full
    inversion when those leave the count open. The oracle inverts every slot
    in full, as the reference's replays do. Larger v0 and T bumps send more
    slots through the second-order and fallback paths; counts must still
    agree exactly (FP64 results to 1e-9)."""
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [H_HI], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(strikes=np.linspace(70, 130, 8), maturity=0.25)
    cfg = mc_config(1 << 14, 100, seed=11, vol_bump=bump, t_bump=bump, num_streams=1024)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL, cfg)
    compare(g, c, FULL, 1e-9)


@pytest.mark.parametrize("scheme", [capi.SCHEME_NV, capi.SCHEME_NCI, capi.SCHEME_EXACT])
@pytest.mark.parametrize("mode", [capi.GREEK_NONE, capi.GREEK_FULL])
def test_schemes_fp64(engine, oracle, scheme, mode):
    h = H1 if scheme == capi.SCHEME_NV else H_HI
    m = make_model(1, [h])
    inst = InstrumentSpec(strikes=[80.0, 100.0, 120.0], maturity=1.0)
    cfg = mc_config(1 << 13, 20, seed=7, num_streams=512)
    g, c = _run_both(engine, oracle, m, [scheme], inst, mode, cfg)
    compare(g, c, FULL if mode == capi.GREEK_FULL else ["price", "price_std_err"], 1e-9)


@pytest.mark.parametrize("jump", ["none", "merton", "kou"])
@pytest.mark.parametrize("nf", [0, 2])
def test_factors_and_jumps_fp64(engine, oracle, jump, nf):
    jk, jp = {"none": (capi.JUMP_NONE, (0, 0, 0, 0)), "merton": merton(0.5, -0.1, 0.15),
              "kou": kou(1.0, 0.4, 10.0, 5.0)}[jump]
    if nf == 0:
        m = make_model(0, vol=0.2, jump_kind=jk, jump=jp)
        schemes = []
    else:
        m = make_model(2, [H1, H2], jump_kind=jk, jump=jp)
        schemes = [capi.SCHEME_NV, capi.SCHEME_NCI]
    inst = InstrumentSpec(strikes=[90.0, 100.0, 110.0], maturity=1.0)
    cfg = mc_config(1 << 13, 16, seed=3, num_streams=512)
    g, c = _run_both(engine, oracle, m, schemes, inst, capi.GREEK_FULL, cfg)
    compare(g, c, FULL, 1e-9)


@pytest.mark.parametrize("kind", ["geo", "arith", "lookback_fixed", "lookback_float",
                                  "barrier_uo", "barrier_di_rebate"])
@pytest.mark.parametrize("jump", [False, True])
def test_path_dependent_fp64(engine, oracle, kind, jump):
    jk, jp = merton(0.5, -0.1, 0.15) if jump else (capi.JUMP_NONE, (0, 0, 0, 0))
    m = make_model(1, [H1], jump_kind=jk, jump=jp)
    spec = dict(strikes=[95.0, 105.0], maturity=1.0)
    if kind == "geo":
        inst = InstrumentSpec(kind=capi.INSTR_GEO_ASIAN, **spec)
    elif kind == "arith":
        inst = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, **spec)
    elif kind == "lookback_fixed":
        inst = InstrumentSpec(kind=capi.INSTR_LOOKBACK, **spec)
    elif kind == "lookback_float":
        inst = InstrumentSpec(kind=capi.INSTR_LOOKBACK, fixed_strike=False, **spec)
    elif kind == "barrier_uo":
        inst = InstrumentSpec(kind=capi.INSTR_BARRIER, barrier=125.0, **spec)
    else:
        inst = InstrumentSpec(kind=capi.INSTR_BARRIER, barrier=85.0, barrier_dir=capi.BARRIER_DOWN,
                              barrier_action=capi.BARRIER_IN, rebate_timing=capi.REBATE_AT_HIT,
                              rebate=2.0, rebate_rate=0.02, **spec)
    cfg = mc_config(1 << 12, 24, seed=11, num_streams=256)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NV], inst, capi.GREEK_FULL, cfg)
    compare(g, c, FULL, 1e-9)


def test_batch_matches_single_calls(engine):
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [H_HI], jump_kind=jk, jump=jp)
    strikes = np.linspace(70, 130, 8)
    insts = [InstrumentSpec(strikes=strikes, maturity=T) for T in (0.1, 0.5, 1.0)]
    cfgs = [mc_config(1 << 12, int(round(100 * T)), seed=42, num_streams=512) for T in (0.1, 0.5, 1.0)]
    jobs = [(m, [capi.SCHEME_NCI], i, capi.GREEK_FULL, c, 100.0, 0.02, 0.0) for i, c in zip(insts, cfgs)]
    batch = engine.mc_batch(jobs)
    for (mm, sch, i, mode, c, S0, r, q), b in zip(jobs, batch):
        single = engine.mc(mm, sch, i, mode, c, S0, r, q)
        for f in FULL:
            assert np.array_equal(np.asarray(single[f]), np.asarray(b[f])), f


def test_mlmc_fp64_matches_oracle(engine, oracle):
    m = make_model(1, [H1])
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0, maturity=1.0)
    cfg = mlmc_config(epsilon=0.05, L_min=2, L_max=5, N0=2000, base_steps=4, num_streams=1024)
    g = engine.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)
    c = oracle.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)
    assert g["n_levels"] == c["n_levels"]
    assert np.array_equal(g["paths_per_level"], c["paths_per_level"])
    assert rel_err(g["price"], c["price"]) < 1e-9
    assert rel_err(g["mean_diffs"], c["mean_diffs"], 1e-10) < 1e-8


def test_mlmc_graph_replays_are_deterministic(engine):
    """Each level runs as a captured HIP graph replayed on every visit of the
    Giles loop: two runs in one context (graphs rebuilt, stream states
    re-seeded) give identical results, in FP32 and FP64."""
    m = make_model(1, [H1])
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0, maturity=1.0)
    for prec in (capi.FP32, capi.FP64):
        cfg = mlmc_config(epsilon=0.02, L_min=3, L_max=6, N0=4000, base_steps=4, precision=prec)
        a = engine.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)
        b = engine.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)
        assert a["n_levels"] == b["n_levels"]
        assert np.array_equal(a["paths_per_level"], b["paths_per_level"])
        assert a["price"] == b["price"] and a["price_std_err"] == b["price_std_err"]


def test_batched_double_heston_sets(engine, oracle):
    """Parameter sets of the batched Double-Heston workload (NV or Exact per
    factor by the Feller-type test), an arithmetic Asian and a down-and-out
    call per set on identical draws: one batch equals the single calls bit
    for bit, and FP64 results equal the oracle on identical draws."""
    import bench
    sets = bench.double_heston_sets(6, log2_paths=(10, 12))
    asian = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=1.0)
    dao = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=80.0, maturity=1.0,
                         barrier_dir=capi.BARRIER_DOWN, barrier_action=capi.BARRIER_OUT)
    jobs = []
    for hs, schemes, paths in sets:
        m = make_model(2, hs)
        cfg = mc_config(paths, 32, seed=42, precision=capi.FP64, num_streams=256)
        jobs += [(m, schemes, asian, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0),
                 (m, schemes, dao, capi.GREEK_NONE, cfg, 100.0, 0.02, 0.0)]
    assert {s for _h, sch, _p in sets for s in sch} == {capi.SCHEME_NV, capi.SCHEME_EXACT}
    batch = engine.mc_batch(jobs)
    for j, b in zip(jobs, batch):
        single = engine.mc(*j)
        assert np.array_equal(single["price"], b["price"])
        assert np.array_equal(single["price_std_err"], b["price_std_err"])
    for j, b in list(zip(jobs, batch))[:4]:
        c = oracle.mc(*j)
        assert rel_err(b["price"], c["price"]) < 1e-9


@pytest.mark.parametrize("prec,mode,nf", [(capi.FP32, capi.GREEK_NONE, 2), (capi.FP64, capi.GREEK_FULL, 1)])
def test_fused_asian_barrier_equal_separate_calls(engine, prec, mode, nf):
    """An arithmetic Asian and a barrier option on identical paths (same
    model, config and seed -- the reference's two calls with one seed) are
    priced in one batch from one simulation: the barrier job's kernel keeps
    the Asian tracker's state too and the Asian is accumulated from its
    staged samples. Results must equal the two separate calls bit for bit."""
    if nf == 2:
        m = make_model(2, [Heston(2.0, 0.04, 0.3, -0.7, 0.04), Heston(1.0, 0.03, 0.9, -0.5, 0.02)])
        sch = [capi.SCHEME_NV, capi.SCHEME_EXACT]
    else:
        m, sch = make_model(1, [H1]), [capi.SCHEME_NV]
    asian = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[95.0, 100.0, 105.0], maturity=1.0)
    dao = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=80.0, maturity=1.0,
                         barrier_dir=capi.BARRIER_DOWN, barrier_action=capi.BARRIER_OUT)
    cfg = mc_config(1 << 15, 126, seed=42, precision=prec, num_streams=2048)
    other = mc_config(1 << 15, 126, seed=43, precision=prec, num_streams=2048)
    fused = engine.mc_batch([(m, sch, dao, mode, cfg, 100.0, 0.02, 0.0),
                             (m, sch, asian, mode, cfg, 100.0, 0.02, 0.0),
                             (m, sch, asian, mode, other, 100.0, 0.02, 0.0)])
    sep = [engine.mc(m, sch, dao, mode, cfg, 100.0, 0.02, 0.0),
           engine.mc(m, sch, asian, mode, cfg, 100.0, 0.02, 0.0),
           engine.mc(m, sch, asian, mode, other, 100.0, 0.02, 0.0)]
    fields = FULL if mode == capi.GREEK_FULL else ["price", "price_std_err"]
    for f_res, s_res in zip(fused, sep):
        for f in fields:
            assert np.array_equal(np.asarray(f_res[f]), np.asarray(s_res[f])), f


@pytest.mark.parametrize("jump", ["merton", "kou"])
def test_jump_means_from_ten(engine, oracle, jump):
    """Jump means from 10 on: Boost's poisson_distribution switches to its
    PTRD rejection sampler there; the engine and the oracle invert the same
    single uniform exactly instead (Giles' method). On identical draws the
    GPU equals the oracle (FP64, Full strip with its T-bumped means on either
    side of 10 for T = 2, lambda 5: means 10 +- 5e-4), and the price agrees
    with the Gil-Pelaez pricer."""
    jk, jp = merton(5.0, -0.05, 0.1) if jump == "merton" else kou(5.0, 0.4, 10.0, 5.0)
    m = make_model(1, [H1], jump_kind=jk, jump=jp)
    inst = InstrumentSpec(strikes=[80.0, 100.0, 120.0], maturity=2.0)    # European: mean 5 x 2
    cfg = mc_config(1 << 12, 50, seed=5, precision=capi.FP64, num_streams=1024)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NV], inst, capi.GREEK_FULL, cfg)
    compare(g, c, FULL, 1e-9)
    big = engine.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE,
                    mc_config(1 << 20, 200, seed=7, precision=capi.FP64), 100.0, 0.02, 0.0)
    f = engine.fourier(m, inst, capi.GREEK_NONE, 100.0, 0.02, 0.0)
    z = np.abs(np.asarray(big["price"]) - np.asarray(f["price"])) / np.asarray(big["price_std_err"])
    assert z.max() < 4.0, z
    # an asian (log_jump on every step) with mean 5 x 2.5 per step span
    asian = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=5.0)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NV], asian, capi.GREEK_NONE,
                     mc_config(1 << 10, 2, seed=5, precision=capi.FP64, num_streams=256))
    compare(g, c, ["price", "price_std_err"], 1e-9)


def test_nci_fine_steps_agree_with_fourier(engine):
    """SchemeNCI at 1,600 steps per year against the Gil-Pelaez prices (Bates,
    sigma 1, T 0.5, 64 strikes, 2^23 paths). Interpolating the table's first
    grid interval -- the reference's rule -- lowers every step's variance
    draw slightly, and the mean-reverting variance adds that up to a drift
    that grows as the step shrinks. At this size the reference's table
    semantics (MYQL_FLAG_REFERENCE_NCI_ROWS) put the strikes at mean z
    +3.2..+3.8 above the Fourier prices (max |z| 6.7, seeds 7 and 42); the
    engine's default inverts that interval exactly and gives -0.02..+0.57
    (max |z| 1.4). The strikes share paths, so their mean z is tested, with
    every |z| < 4."""
    jk, jp = merton(0.5, -0.1, 0.15)
    m = make_model(1, [H_HI], jump_kind=jk, jump=jp)
    strikes = list(np.linspace(70.0, 130.0, 64))
    inst = InstrumentSpec(strikes=strikes, maturity=0.5)
    print("nci fine steps: 2^23 paths x 800 steps", flush=True)
    mc = engine.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_NONE,
                   mc_config(1 << 23, 800, seed=42, precision=capi.FP32, num_streams=1 << 18), 100.0, 0.02, 0.0)
    fo = engine.fourier(m, inst, capi.GREEK_NONE, 100.0, 0.02, 0.0)
    ok = fo["price"] >= 0.1
    z = ((mc["price"] - fo["price"]) / mc["price_std_err"])[ok]
    assert abs(z.mean()) < 2.0 and np.abs(z).max() < 4.0, (z.mean(), np.abs(z).max())


def test_giles_and_acklam_against_reference_numerics(engine):
    """Device Giles counts equal the reference's own Numerics::poissinvcdf
    (compiled from math/Numerics.hpp; tests/golden/numerics_ref.json) on all
    2,920 (lambda, U) pairs; the device Acklam quantile is within 1e-12
    relative of the reference's (FMA contraction and the device's log; 2e-13
    measured, at quantiles near 0 where the central branch cancels)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "numerics_ref.json")) as f:
        d = json.load(f)
    n = engine.poissinv(np.array(d["lam"]), np.array(d["u"]))
    assert np.array_equal(n, np.array(d["n"]))
    z64, _ = engine.normcdfinv(np.array(d["p"]))
    ref = np.array(d["z"])
    assert np.max(np.abs(z64 - ref) / np.abs(ref)) < 1e-12


@pytest.mark.parametrize("kind", ["geo_in_flight", "arith_in_flight", "lookback_float_in_flight",
                                  "barrier_di_rebate_at_maturity"])
def test_in_flight_and_rebate_variants_fp64(engine, oracle, kind):
    """In-flight Asians (observed averages over an elapsed period), an
    in-flight floating lookback (observed extremes) and a down-and-in put with
    a rebate paid at maturity: FP64 engine against the oracle on identical
    draws (the oracle's trackers and payoffs equal the reference's own on
    fixed paths, test_oracle.py)."""
    m = make_model(1, [H1])
    spec = dict(strikes=[95.0, 105.0], maturity=1.0)
    if kind == "geo_in_flight":
        inst = InstrumentSpec(kind=capi.INSTR_GEO_ASIAN, past_avg_log=float(np.log(95.0)), t_elapsed=0.25, **spec)
    elif kind == "arith_in_flight":
        inst = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, past_avg_S=97.0, past_avg_log=float(np.log(97.0)),
                              t_elapsed=0.5, **spec)
    elif kind == "lookback_float_in_flight":
        inst = InstrumentSpec(kind=capi.INSTR_LOOKBACK, fixed_strike=False, hist_min_S=90.0, hist_max_S=115.0,
                              **spec)
    else:
        inst = InstrumentSpec(kind=capi.INSTR_BARRIER, barrier=90.0, barrier_dir=capi.BARRIER_DOWN,
                              barrier_action=capi.BARRIER_IN, option_type=capi.PUT,
                              rebate_timing=capi.REBATE_AT_MATURITY, rebate=1.5, rebate_rate=0.03, **spec)
    cfg = mc_config(1 << 12, 24, seed=11, num_streams=256)
    g, c = _run_both(engine, oracle, m, [capi.SCHEME_NV], inst, capi.GREEK_FULL, cfg)
    compare(g, c, FULL, 1e-9)


def test_device_trackers_and_payoffs_match_reference_fixture(engine):
    """The device's trackers and calculate_to_buffer (FP64) on the fixed
    log-price paths of tests/golden/instruments_ref.json, the reference's own
    trackers, payoffs and option headers compiled from its headers: base,
    spot-up and spot-down payoffs of 21 instruments on 8 paths within 1e-13
    of max(1, |value|) (FMA contraction on the device; the oracle equals the
    fixture bit for bit)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "instruments_ref.json")) as f:
        fx = json.load(f)
    kind = {"european": capi.INSTR_EUROPEAN, "geo_asian": capi.INSTR_GEO_ASIAN,
            "arith_asian": capi.INSTR_ARITH_ASIAN, "lookback": capi.INSTR_LOOKBACK, "barrier": capi.INSTR_BARRIER}
    pay = {"vanilla": capi.PAYOFF_VANILLA, "cash": capi.PAYOFF_CASH, "asset": capi.PAYOFF_ASSET}
    reb = {"none": capi.REBATE_NONE, "at_hit": capi.REBATE_AT_HIT, "at_maturity": capi.REBATE_AT_MATURITY}
    worst = 0.0
    for c in fx["cases"]:
        d = c["spec"]
        inst = InstrumentSpec(
            kind=kind[d["kind"]], payoff=pay[d.get("payoff", "vanilla")],
            option_type=capi.CALL if d["type"] == "call" else capi.PUT, strikes=d["strikes"],
            maturity=d["maturity"], is_strip=len(d["strikes"]) > 1, fixed_strike=d.get("fixed_strike", True),
            barrier_dir=capi.BARRIER_UP if d.get("dir", "up") == "up" else capi.BARRIER_DOWN,
            barrier_action=capi.BARRIER_OUT if d.get("action", "out") == "out" else capi.BARRIER_IN,
            rebate_timing=reb[d.get("rebate_timing", "none")], barrier=d.get("barrier", 0.0),
            rebate=d.get("rebate", 0.0), rebate_rate=d.get("rebate_rate", 0.0), payoff_eps=d.get("payoff_eps", -1.0),
            past_avg_log=d.get("past_avg_log", 0.0), past_avg_S=d.get("past_avg_S", 0.0),
            t_elapsed=d.get("t_elapsed", 0.0), hist_min_S=d.get("hist_min_S", 0.0),
            hist_max_S=d.get("hist_max_S", 0.0))
        b, u, dn = engine.path_payoffs(inst, capi.GREEK_ESSENTIAL, fx["S0"], fx["fd_bump"], [c["logS"]])
        for got, ref in ((b[0], c["base"]), (u[0], c["up"]), (dn[0], c["dn"])):
            ref = np.asarray(ref)
            err = float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))))
            worst = max(worst, err)
            assert err < 1e-13, (c["name"], c["path"], got, ref)
    print("worst relative difference", worst)
