"""Headline benchmark: the Bates strip.

Workload (one "step"): one pricing of the whole strip through
myql_mc_calculate_batch -- 64 strikes uniform in [70, 130] x 8 maturities
{0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0}, each maturity a separate
MonteCarloPricer::calculate call (2^24 paths on 262,144 xoshiro streams, 100 steps per year, seed 42,
GreekMode::Full with common random numbers) under Bates with SchemeNCI:
kappa 2, theta 0.04, sigma 1.0 (Feller violated), rho -0.7, v0 0.04, Merton
jumps lambda 0.5, mu -0.1, delta 0.15; S0 100, r 0.02, q 0. Paths are FP32,
sums FP64. Inputs are synthetic (no market data).

Metric: path-steps/s = (sum over maturities of paths x steps x 7 Greek
variants) / wall time, the 7 variants being the reference algorithm's
base, v0 up/down, T up/down and r up/down replays. Also reported: ms to a
1e-4 relative standard error (time x (worst se/price over options with
price >= 0.1 / 1e-4)^2) for price-only plain FP32, plain FP64 and the learned
martingale control variate (training included) on the strip's 512 options,
the same for the Full-Greek runs, and the control variate on a Bates-Kou
SchemeNV strip; the CPU baseline; the roofline of the path kernel.

The strip kernel's launch shape (streams per work unit, strike-loop unroll,
grid factor) is autotuned on an untimed first call (MYQL_FLAG_AUTOTUNE) and
reported under config.launch_shape.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--paths-log2 24] [--no-cpu] [--no-fp64]
                [--no-cv] [--no-mlmc] [--no-batched] [--no-autotune] [--shape B,U,G[,W4]] [--copy-probe]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Bates strip: path-steps/s; ms to 1e-4 rel std-err, plain vs learned CV"
MATURITIES = [0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0]
STRIKES = np.linspace(70.0, 130.0, 64)
S0, R, Q = 100.0, 0.02, 0.0
VARIANTS = 7
# Algorithmic floating-point operations per path-step of this workload: the
# reference algorithm as written (each of the 7 variants replays its path; an
# FMA counts 2; add, mul, div, sqrt, exp, log, erfc 1 each; stepping, Poisson
# and chi-squared inversion, jumps, payoffs, Greek estimators and sums; NCI
# table set-up excluded), counted by the oracle's op-counting build:
#   python oracle/count_flops.py --paths-log2 16   -> 103.67
FLOP_PER_PATH_STEP = 103.69
# The same count under the reference's NCI rows and exact chi-squared
# fallback above them (MYQL_FLAG_REFERENCE_NCI_ROWS; VolSchemes.hpp:245-255,
# 334-370): the fallback fires on ~4 % of steps and its Halley iterations
# raise the count:
#   python oracle/count_flops.py --paths-log2 16 --reference-nci-rows -> 143.61
FLOP_PER_PATH_STEP_REF_ROWS = 143.61
# The same count for the batched Double-Heston workload (each set priced as
# the reference's two calls, weighted by the sets' path counts) and for the
# MLMC workload (per path-step as the engine counts them: fine plus coarse):
#   python oracle/count_flops.py --workload batched -> 137.35
#   python oracle/count_flops.py --workload mlmc    -> 45.53
FLOP_PER_PATH_STEP_BATCHED = 137.35
FLOP_PER_PATH_STEP_MLMC = 45.53
# MI355X published FP32 vector peak (packed dual-issue FP32; scalar FP32 and
# FP64 vector issue peak at 78.6 TFLOP/s)
FP32_PEAK_TFLOPS = 157.3
# Path variants the engine simulates per sample for this European strip: the
# r +- variants are the base path shifted by +-r_bump*T in log space (only the
# drift depends on r), so 5 of the 7 are stepped
EXECUTED_VARIANTS = 5
# HBM bytes per path-kernel launch (FETCH_SIZE + WRITE_SIZE from rocprofv3,
# scaled by the same counters on a plain copy of known size), by
# --paths-log2 of the measured run; None where not measured
TRAFFIC_BYTES_PER_LAUNCH: dict = {24: 4.14e11}


def model():
    from myquantlib_amd.pricing import Heston, make_model, merton
    jk, jp = merton(0.5, -0.1, 0.15)
    return make_model(1, [Heston(2.0, 0.04, 1.0, -0.7, 0.04)], jump_kind=jk, jump=jp)


# xoshiro streams of every strip call: 64 samples per stream at 2^24 paths,
# 4,096 work units per maturity, so the persistent queue balances to within a
# fraction of a unit (65,536 streams leave ~2 units per wavefront slot and a
# 480-550 ms spread from run to run; 262,144 run at a steady ~450 ms)
STRIP_STREAMS = 1 << 18


def jobs(paths: int, precision: int, layout: int = 0, streams: int = STRIP_STREAMS, mode: int = None,
         seed: int = 42, bates_kou: bool = False, flags: int = 0, payoff: int = 0):
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec, mc_config
    m = model_bates_kou() if bates_kou else model()
    scheme = capi.SCHEME_NV if bates_kou else capi.SCHEME_NCI
    mode = capi.GREEK_FULL if mode is None else mode
    out = []
    for T in MATURITIES:
        inst = InstrumentSpec(payoff=payoff, strikes=STRIKES, maturity=T)
        cfg = mc_config(paths, int(round(100 * T)), seed=seed, precision=precision,
                        stream_layout=layout, num_streams=streams, flags=flags)
        out.append((m, [scheme], inst, mode, cfg, S0, R, Q))
    return out


def model_bates_kou():
    """Bates-Kou with SchemeNV (sigma^2 <= 4 kappa theta): Heston kappa 2,
    theta 0.04, sigma 0.5, rho -0.7, v0 0.04; Kou lambda 1, p 0.4, eta1 10,
    eta2 5 -- the learned control variate's second workload."""
    from myquantlib_amd.pricing import Heston, kou, make_model
    jk, jp = kou(1.0, 0.4, 10.0, 5.0)
    return make_model(1, [Heston(2.0, 0.04, 0.5, -0.7, 0.04)], jump_kind=jk, jump=jp)


def cv_leg(eng, eval_paths: int, bates_kou: bool, payoff: int = 0, points: int = 16):
    """Plain FP32, plain FP64 and learned-CV FP32 on the strip's 512 options
    (vanilla calls, price only; or auto-smoothed cash digitals, which the
    reference smooths in Essential mode, priced in that mode): ms to 1e-4
    relative SE for each. The CV network is trained once (2^20 training
    paths per maturity, seed 7, 2,000 Adam steps of 8,192 rows, lr 1e-3,
    16 control points; the regressor coefficients are fitted on the same
    training paths) and then evaluated on independent paths (seed 42), so
    its time to the target is the training time plus the evaluation time
    scaled to the target: t_train + t_eval (worst se / price / 1e-4)^2."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import cv_config
    mode = capi.GREEK_NONE if payoff == capi.PAYOFF_VANILLA else capi.GREEK_ESSENTIAL
    J = jobs(eval_paths, capi.FP32, mode=mode, bates_kou=bates_kou, payoff=payoff)
    J64 = jobs(eval_paths, capi.FP64, mode=mode, bates_kou=bates_kou, payoff=payoff)
    timed = {}
    for name, jj in (("fp32", J), ("fp64", J64)):
        eng.mc_batch(jj)
        t0 = time.perf_counter()
        r = eng.mc_batch(jj)
        timed[name] = (r, 1e3 * (time.perf_counter() - t0))
    plain, tp = timed["fp32"]
    plain64, tp64 = timed["fp64"]
    cfg = cv_config(train_paths=1 << 20, train_seed=7, steps=2000, batch=8192, lr=1e-3, points=points)
    t0 = time.perf_counter()
    net = eng.cv_train(jobs(1024, capi.FP32, mode=mode, seed=7, bates_kou=bates_kou, payoff=payoff), cfg)
    tt = 1e3 * (time.perf_counter() - t0)
    eng.mc_batch_cv(J, cfg, net)  # warm-up
    t0 = time.perf_counter()
    res = eng.mc_batch_cv(J, cfg, net)
    tc = 1e3 * (time.perf_counter() - t0)
    mp, wp = ms_to_target(plain, tp)
    m64, w64 = ms_to_target(plain64, tp64)
    _, wc = ms_to_target(res, tc)
    mc = tt + tc * (wc / 1e-4) ** 2
    vr = np.concatenate([r["var_ratio"] for r in res])
    # the CV prices of the vanillas against the GPU Gil-Pelaez pricer, in CV
    # standard errors (options with price >= 0.1)
    fz = None
    if payoff == capi.PAYOFF_VANILLA:
        from myquantlib_amd.pricing import InstrumentSpec
        m = model_bates_kou() if bates_kou else model()
        fr = eng.fourier_batch([(m, InstrumentSpec(strikes=STRIKES, maturity=T), capi.GREEK_NONE, S0, R, Q)
                                for T in MATURITIES])
        z = []
        for r, f in zip(res, fr):
            ok = f["price"] >= 0.1
            z.extend(np.abs((r["price"] - f["price"]) / r["price_std_err"])[ok])
        z = np.asarray(z)
        fz = {"max_abs_z": float(z.max()), "frac_within_3se": float(np.mean(z < 3))}
    used = int(sum(int(np.sum(r["used_cv"])) for r in res))
    # the options that set the metric, plain and CV, and their variance ratios
    def worst_option(rs):
        best = (0.0, None)
        for ti, r in enumerate(rs):
            ok = r["price"] >= 0.1
            rel = np.where(ok, r["price_std_err"] / np.maximum(r["price"], 1e-300), 0.0)
            k = int(np.argmax(rel))
            if rel[k] > best[0]:
                best = (float(rel[k]), (MATURITIES[ti], float(STRIKES[k]), float(res[ti]["var_ratio"][k])))
        return best[1]
    label = ("bates-kou NV" if bates_kou else "bates NCI") + (" digitals" if payoff else "")
    print(f"# cv leg ({label}): plain fp32 {tp:.0f} ms, fp64 {tp64:.0f} ms, cv {tc:.0f} ms + train {tt:.0f} ms, "
          f"median variance ratio {np.median(vr):.2f}, ms to 1e-4: plain fp64 {m64:.3g}, cv {mc:.3g} "
          f"({m64 / mc:.2f}x)", file=sys.stderr, flush=True)
    return {"payoff": "cash digital, auto-smoothed (Essential)" if payoff else "vanilla call",
            "eval_paths_per_maturity": eval_paths, "control_points": points,
            "plain_fp32_ms": tp, "plain_fp32_ms_to_1e-4": mp, "plain_worst_rel_se": wp,
            "plain_fp64_ms": tp64, "plain_fp64_ms_to_1e-4": m64,
            "learned_cv_fp32_ms": tc, "train_ms": tt, "train_device_ms": net.train_ms,
            "learned_cv_fp32_ms_to_1e-4": mc, "cv_worst_rel_se": wc,
            "learned_cv_fp32_ms_to_1e-4_training_per_run": (tt + tc) * (wc / 1e-4) ** 2,
            "speedup_vs_plain_fp64": m64 / mc,
            "worst_option_plain": {"T_K_var_ratio": worst_option(plain)},
            "worst_option_cv": {"T_K_var_ratio": worst_option(res)},
            "cv_vs_fourier": fz,
            "variance_ratio_median": float(np.median(vr)), "variance_ratio_min": float(np.min(vr)),
            "variance_ratio_max": float(np.max(vr)),
            "options_using_cv": used, "options": int(vr.size), "train_loss": net.final_loss}


def qmc_leg(eng, paths: int, po32_ms_to, reps: int = 16):
    """Randomized quasi-Monte Carlo on the price-only strip (GreekMode None,
    FP32 paths, the strip's model and maturities): `paths` Sobol' points per
    maturity in `reps` digitally shifted replicates (MYQL_STREAMS_SOBOL;
    3 coordinates per step, 602 at T = 2), against pseudo-random FP32 on
    the same number of paths; ms to 1e-4 relative SE as for the other legs
    (the replicates' standard error for QMC)."""
    from myquantlib_amd import capi
    Jq = jobs(paths, capi.FP32, mode=capi.GREEK_NONE, streams=min(paths // reps, STRIP_STREAMS))
    for j in Jq:
        j[4].stream_layout = capi.STREAMS_SOBOL
        j[4].qmc_replicates = reps
    Jp = jobs(paths, capi.FP32, mode=capi.GREEK_NONE, streams=min(paths, STRIP_STREAMS))
    eng.mc_batch(Jq)
    t0 = time.perf_counter()
    rq = eng.mc_batch(Jq)
    tq = 1e3 * (time.perf_counter() - t0)
    eng.mc_batch(Jp)
    t0 = time.perf_counter()
    rp = eng.mc_batch(Jp)
    tp = 1e3 * (time.perf_counter() - t0)
    mq, wq = ms_to_target(rq, tq)
    mp, wp = ms_to_target(rp, tp)
    ratios = []
    for a, b in zip(rq, rp):
        ok = np.asarray(b["price"]) >= 0.1
        ratios.extend((np.asarray(b["price_std_err"]) / np.asarray(a["price_std_err"]))[ok])
    out = {"paths_per_maturity": paths, "replicates": reps, "qmc_fp32_ms": tq, "plain_fp32_ms": tp,
           "qmc_fp32_ms_to_1e-4": mq, "plain_fp32_ms_to_1e-4": mp, "qmc_worst_rel_se": wq, "plain_worst_rel_se": wp,
           "se_ratio_median": float(np.median(ratios)), "se_ratio_min": float(np.min(ratios)),
           "speedup_vs_plain_fp32": mp / mq, "full_strip_plain_fp32_ms_to_1e-4": po32_ms_to}
    print(f"# qmc leg: {tq:.0f} ms (plain {tp:.0f} ms), median SE ratio {out['se_ratio_median']:.2f}, "
          f"ms to 1e-4: qmc {mq:.3g}, plain {mp:.3g} ({mp / mq:.2f}x)", file=sys.stderr, flush=True)
    return out


def path_steps(paths: int) -> int:
    return sum(paths * int(round(100 * T)) * VARIANTS for T in MATURITIES)


def ms_to_target(results, t_run_ms: float, target: float = 1e-4):
    worst = 0.0
    for res in results:
        p, se = res["price"], res["price_std_err"]
        ok = p >= 0.1
        if np.any(ok):
            worst = max(worst, float(np.max(se[ok] / p[ok])))
    return t_run_ms * (worst / target) ** 2, worst


def host_cpu():
    """(nproc, logical CPUs, CPU model name) of this host. `nproc` counts the
    cores this process may use and honours OMP_NUM_THREADS, as an OpenMP run
    with the default team size would."""
    import subprocess
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except (OSError, ValueError, subprocess.CalledProcessError):
        nproc = len(os.sched_getaffinity(0))
    name = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return nproc, os.cpu_count() or nproc, name


def cpu_baseline(target_s: float = 8.0):
    """The oracle compiled with the reference's flags, OpenMP over host cores
    in the reference's structure (per-thread stream seeded seed + tid*999983,
    static partition). Sample: the T = 1 maturity of the strip (100 steps,
    Full CRN Greeks, 64 strikes), path count sized to ~target_s seconds; one
    warm-up run of that size, then the median of 3 timed runs. Threads =
    `nproc` (OMP_NUM_THREADS=$(nproc), OMP_PROC_BIND=close)."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec, mc_config
    from oracle.oracle import Oracle
    nproc, logical, cpu_name = host_cpu()
    cores = int(os.environ.get("MYQL_CPU_THREADS", nproc))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    o = Oracle(fast=True)
    m = model()
    inst = InstrumentSpec(strikes=STRIKES, maturity=1.0)

    def run(paths):
        cfg = mc_config(paths, 100, seed=42, stream_layout=capi.STREAMS_SEEDED, num_streams=cores)
        t0 = time.perf_counter()
        o.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL, cfg, S0, R, Q, nthreads=cores)
        return time.perf_counter() - t0

    run(cores * 16)  # warm-up
    # two pilots separate the per-call set-up (the NCI table build, which the
    # reference also repeats in every calculate call) from the per-path cost
    p1, p2 = cores * 256, cores * 2048
    t1, t2 = run(p1), run(p2)
    per_path = max((t2 - t1) / (p2 - p1), 1e-9)
    fixed = max(t1 - p1 * per_path, 0.0)
    paths = int(max(p2, min(1 << 23, (target_s - fixed) / per_path)))
    run(paths)  # warm-up at the timed size
    times = sorted(run(paths) for _ in range(3))
    t = times[1]
    rate = paths * 100 * VARIANTS / t
    print(f"# cpu baseline: {cores} threads on {cpu_name}, runs {[round(x, 2) for x in times]} s, "
          f"median {rate:.3e} path-steps/s", file=sys.stderr, flush=True)
    return {"value": rate, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "nproc": nproc, "logical_cpus": logical, "cpu_model": cpu_name,
            "runs_s": times, "statistic": "median of 3 after 1 warm-up",
            "sample": f"T=1 maturity of the strip: {paths} paths x 100 steps x 7 CRN variants, "
                      f"64 strikes, Full Greeks, FP64, median {t:.1f} s of 3 on {cores} threads "
                      f"({cpu_name})"}


CV_LEGS = (("bates_strip", False, False), ("bates_kou_nv_strip", True, False),
           ("bates_kou_nv_smoothed_digitals", True, True))


def cv_nets(eng, path: str, save: bool):
    """The three control-variate legs split in two processes, so that the
    pricing kernels can be profiled apart from the training: save=True
    trains each leg's network (as cv_leg does) and writes the myql_cv_net
    struct to path.<i>; save=False reads them back and times the pricing
    call alone (2^22 paths per maturity, after one warm-up call)."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import cv_config
    cfg = cv_config(train_paths=1 << 20, train_seed=7, steps=2000, batch=8192, lr=1e-3, points=16)
    out = {}
    for i, (name, kou, dig) in enumerate(CV_LEGS):
        payoff = capi.PAYOFF_CASH if dig else capi.PAYOFF_VANILLA
        mode = capi.GREEK_ESSENTIAL if dig else capi.GREEK_NONE
        fn = f"{path}.{i}"
        if save:
            net = eng.cv_train(jobs(1024, capi.FP32, mode=mode, seed=7, bates_kou=kou, payoff=payoff), cfg)
            with open(fn, "wb") as f:
                f.write(bytes(net))
            out[name] = {"train_ms": net.train_ms}
        else:
            with open(fn, "rb") as f:
                net = capi.CVNet.from_buffer_copy(f.read())
            J = jobs(1 << 22, capi.FP32, mode=mode, bates_kou=kou, payoff=payoff)
            eng.mc_batch_cv(J, cfg, net)
            t0 = time.perf_counter()
            res = eng.mc_batch_cv(J, cfg, net)
            out[name] = {"cv_ms": 1e3 * (time.perf_counter() - t0),
                         "variance_ratio_median": float(np.median(np.concatenate([r["var_ratio"] for r in res])))}
        print(f"# cv nets {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def cv_jump_tables(eng):
    """The control variate's terminal-jump tables of the Bates-Kou strip alone
    (2,048 log-spot points x 64 strikes per maturity; vanilla calls, and
    cash digitals smoothed with the strip's automatic bandwidth at T = 1),
    built through myql_debug_cv_jump_expectation, which builds the table
    the pricing call builds and reads it back at one spot; wall ms per strip
    of 8 tables, after one warm-up strip."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec
    m = model_bates_kou()
    out = {}
    for name, payoff, mode, eps in (("vanilla", capi.PAYOFF_VANILLA, capi.GREEK_NONE, -1.0),
                                    ("smoothed_cash_digital", capi.PAYOFF_CASH, capi.GREEK_ESSENTIAL, 2.0)):
        ins = [InstrumentSpec(payoff=payoff, strikes=STRIKES, maturity=T, payoff_eps=eps) for T in MATURITIES]
        for rep in range(2):
            t0 = time.perf_counter()
            for i in ins:
                assert eng.cv_jump_expectation(m, i, mode, S0, np.array([S0])) is not None
            ms = 1e3 * (time.perf_counter() - t0)
        out[name + "_strip_ms"] = ms
        print(f"# cv jump tables ({name}): {ms:.1f} ms per strip of {len(ins)}", file=sys.stderr, flush=True)
    return out


def copy_probe(nbytes: int = 1 << 30, reps: int = 3):
    """A plain device-to-device copy of known size: each launch reads and
    writes nbytes, so FETCH_SIZE and WRITE_SIZE of these dispatches give the
    counters' scale (expected nbytes / 1024 KiB each)."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for i in range(reps):
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy {i}: {nbytes} bytes read, {nbytes} bytes written", file=sys.stderr,
              flush=True)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def _xoshiro_uniforms(seed: int):
    """uniform(0,1) draws of xoshiro256++ seeded by splitmix64 (the engine's
    generator), for the parameter sets of the batched workload."""
    M64 = (1 << 64) - 1
    x, st = seed, []
    for _ in range(4):
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        st.append(z ^ (z >> 31))

    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & M64

    while True:
        out = (rotl((st[0] + st[3]) & M64, 23) + st[0]) & M64
        t = (st[1] << 17) & M64
        st[2] ^= st[0]
        st[3] ^= st[1]
        st[1] ^= st[2]
        st[0] ^= st[3]
        st[2] ^= t
        st[3] = rotl(st[3], 45)
        u = out * 2.0 ** -64
        if u < 1.0:
            yield u


def double_heston_sets(n_sets: int = 1024, seed: int = 42, log2_paths=(14, 20)):
    """Parameter sets of the batched workload, drawn in this order per set:
    for factor 1 then factor 2: kappa ~ U[0.5, 5], theta ~ U[0.01, 0.09],
    sigma ~ U[0.1, 1], rho ~ U[-0.9, 0], v0 ~ U[0.01, 0.09]; then the path
    count 2^u, u ~ U[14, 20] (rounded). Scheme per factor: SchemeNV where
    sigma^2 <= 4 kappa theta, else SchemeExact."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import Heston
    g = _xoshiro_uniforms(seed)
    out = []
    for _ in range(n_sets):
        hs, schemes = [], []
        for _f in range(2):
            k = 0.5 + 4.5 * next(g)
            th = 0.01 + 0.08 * next(g)
            sg = 0.1 + 0.9 * next(g)
            rh = -0.9 * next(g)
            v0 = 0.01 + 0.08 * next(g)
            hs.append(Heston(k, th, sg, rh, v0))
            schemes.append(capi.SCHEME_NV if sg * sg <= 4.0 * k * th else capi.SCHEME_EXACT)
        lo, hi = log2_paths
        paths = int(round(2.0 ** (lo + (hi - lo) * next(g))))
        out.append((hs, schemes, paths))
    return out


def batched_leg(eng, n_sets: int = 1024, steps: int = 252, flags: int = 0):
    """1,024 Double-Heston parameter sets, each pricing an arithmetic Asian
    call (K 100) and a down-and-out call (K 100, B 80), T 1, 252 steps, price
    only, FP32 paths; the 2,048 pricing calls go to the device as one batch
    (one persistent launch per kernel instantiation). All sets use seed 42,
    so a set's two instruments see identical draws, as two reference calls
    with the same seed would; the engine detects that and simulates each
    set's paths once, pricing the Asian from the barrier job's staged samples
    (path_steps counts both calls, as the reference runs them). With flags =
    FLAG_AUTOTUNE the warm-up call (the first 128 sets) also picks the
    kernel's launch shape."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec, make_model, mc_config
    sets = double_heston_sets(n_sets)
    asian = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=1.0)
    dao = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=80.0, maturity=1.0,
                         barrier_dir=capi.BARRIER_DOWN, barrier_action=capi.BARRIER_OUT)
    J = []
    for hs, schemes, paths in sets:
        m = make_model(2, hs)
        # 16,384 streams per call (paths >= 2^14): one shared stream table
        cfg = mc_config(paths, steps, seed=42, precision=capi.FP32, num_streams=1 << 14, flags=flags)
        J.append((m, schemes, asian, capi.GREEK_NONE, cfg, S0, R, Q))
        J.append((m, schemes, dao, capi.GREEK_NONE, cfg, S0, R, Q))
    eng.mc_batch(J[:256])  # warm-up: code objects, stream states (and launch shape)
    t0 = time.perf_counter()
    res = eng.mc_batch(J)
    wall = 1e3 * (time.perf_counter() - t0)
    ps = sum(2 * paths * steps for _h, _s, paths in sets)
    n_exact = sum(s.count(capi.SCHEME_EXACT) for _h, s, _p in sets)
    print(f"# batched double heston: {len(J)} calls, {wall:.0f} ms, {ps:.3e} path-steps",
          file=sys.stderr, flush=True)
    dev_s = res[0]["device_ms"] / 1e3
    achieved = ps * FLOP_PER_PATH_STEP_BATCHED / dev_s / 1e12
    return {"workload": f"{n_sets} Double-Heston sets x (arithmetic Asian call K 100, down-and-out "
                        f"call K 100 B 80), T 1, {steps} steps, 2^U[14,20] paths per set, price only, "
                        "FP32 paths, one batch",
            "calls": len(J), "paths": int(sum(p for _h, _s, p in sets)), "path_steps": ps,
            "simulated_path_steps": ps // 2,
            "ms": wall, "device_ms": res[0]["device_ms"], "path_steps_per_s": ps / (wall / 1e3),
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "flop_per_path_step": FLOP_PER_PATH_STEP_BATCHED,
                         "basis": "the reference's two calls per set (path_steps)"},
            "factors_exact": n_exact, "factors_nv": 2 * n_sets - n_exact,
            "price_range": [float(min(r["price"][0] for r in res)), float(max(r["price"][0] for r in res))]}


def batched_by_scheme(eng, n_sets: int = 1024, steps: int = 252):
    """The batched workload's sets split by how many factors take
    SchemeExact (0, 1, 2), each class priced as one batch at a fixed launch
    shape (64-stream units, strike loop unrolled by 4, grid 4x): device time
    per 1e9 path-steps of each class."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec, make_model, mc_config
    eng.set_launch_shape(capi.FP32, 2, 1, True, 64, 4, 4)
    sets = double_heston_sets(n_sets)
    asian = InstrumentSpec(kind=capi.INSTR_ARITH_ASIAN, strikes=[100.0], maturity=1.0)
    dao = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=80.0, maturity=1.0,
                         barrier_dir=capi.BARRIER_DOWN, barrier_action=capi.BARRIER_OUT)

    def jobs(sel):
        J = []
        for hs, schemes, paths in sel:
            m = make_model(2, hs)
            cfg = mc_config(paths, steps, seed=42, precision=capi.FP32, num_streams=1 << 14)
            J += [(m, schemes, asian, capi.GREEK_NONE, cfg, S0, R, Q), (m, schemes, dao, capi.GREEK_NONE, cfg, S0, R, Q)]
        return J

    eng.mc_batch(jobs(sets[:64]))
    out = {}
    for n_exact in (0, 1, 2):
        sel = [x for x in sets if x[1].count(capi.SCHEME_EXACT) == n_exact]
        res = eng.mc_batch(jobs(sel))
        ps = sum(p * steps for _h, _s, p in sel)   # simulated once per set
        out[f"exact_factors_{n_exact}"] = {"sets": len(sel), "path_steps": ps, "device_ms": res[0]["device_ms"],
                                           "ms_per_1e9_path_steps": res[0]["device_ms"] / (ps / 1e9)}
        print(f"# batched, sets with {n_exact} SchemeExact factor(s): {len(sel)} sets, {ps:.3e} path-steps, "
              f"{res[0]['device_ms']:.1f} ms device", file=sys.stderr, flush=True)
    return out


def reference_rows_leg(eng, paths: int, flags: int, steps: int = 3, fp64: bool = True):
    """The headline strip under the reference's NCI semantics
    (MYQL_FLAG_REFERENCE_NCI_ROWS: the reference's row count from
    v_max = max(v0, 3 theta), interpolation of every u < 0.999 including the
    first grid interval, exact chi-squared inversion otherwise;
    VolSchemes.hpp:245-255, 334-370), FP32 and FP64 paths, same seed and
    streams: wall ms per strip, path-steps/s and the path kernel's roofline
    at the reference algorithm's op count."""
    from myquantlib_amd import capi
    out = {}
    total = path_steps(paths)
    for name, prec, n in (("fp32", capi.FP32, steps), ("fp64", capi.FP64, 1 if fp64 else 0)):
        if n == 0:
            continue
        J = jobs(paths, prec, flags=flags | capi.FLAG_REFERENCE_NCI_ROWS)
        eng.mc_batch(J)  # warm-up (launch shape, tables)
        dev = []
        t0 = time.perf_counter()
        for _ in range(n):
            res = eng.mc_batch(J)
            dev.append(res[0]["device_ms"])
        ms = 1e3 * (time.perf_counter() - t0) / n
        kern_s = float(np.mean(dev)) / 1e3
        ach = total * FLOP_PER_PATH_STEP_REF_ROWS / kern_s / 1e12
        _, worst = ms_to_target(res, ms)
        print(f"# reference NCI rows, {name} strip: {ms:.1f} ms/step, device {np.mean(dev):.1f} ms",
              file=sys.stderr, flush=True)
        out[name] = {"ms_per_step": ms, "device_ms": float(np.mean(dev)), "path_steps_per_s": total / (ms / 1e3),
                     "worst_rel_se": worst,
                     "roofline": {"bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach / FP32_PEAK_TFLOPS,
                                  "flop_per_path_step": FLOP_PER_PATH_STEP_REF_ROWS}}
    return out


def fp32_vs_fp64(r32, r64):
    """Error of the FP32-path strip against the FP64-path strip on identical
    draws (same seed and streams; sums are FP64 in both), per result field:
    max |fp32 - fp64| / max(|fp64|, the field's standard error), over options
    with price >= 0.1 -- the FP32 error in units of the result or of its
    Monte Carlo noise, whichever is larger."""
    out = {}
    for f in ("price", "delta", "gamma", "vega", "theta", "rho", "vanna", "charm"):
        worst = 0.0
        for a, b in zip(r32, r64):
            ok = b["price"] >= 0.1
            x, y = np.asarray(a[f], float), np.asarray(b[f], float)
            se = np.asarray(b[f + "_std_err"], float)
            if x.ndim == 2:
                x, y, se = x[0], y[0], se[0]
            d = np.abs(x - y)[ok] / np.maximum(np.abs(y), se)[ok]
            if d.size:
                worst = max(worst, float(d.max()))
        out[f] = worst
    return out


def fourier_check(eng, mc_results):
    """The strip's 512 options with Full Greeks by the GPU Gil-Pelaez pricer
    (one launch), and the FP32 Monte Carlo strip against it: z = (MC - Fourier)
    / MC standard error per option with price >= 0.1 (the mean of the signed
    z shows a bias common to the strip)."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec
    FJ = [(model(), InstrumentSpec(strikes=STRIKES, maturity=T), capi.GREEK_FULL, S0, R, Q) for T in MATURITIES]
    eng.fourier_batch(FJ)
    t0 = time.perf_counter()
    fr = eng.fourier_batch(FJ)
    ms = 1e3 * (time.perf_counter() - t0)
    z = []
    for mc, fo in zip(mc_results, fr):
        ok = fo["price"] >= 0.1
        z.extend(((mc["price"] - fo["price"]) / mc["price_std_err"])[ok])
    zs = np.asarray(z)
    z = np.abs(zs)
    print(f"# fourier strip (512 options, Full): {ms:.2f} ms; MC fp32 vs Fourier max |z| {z.max():.2f}, "
          f"{np.mean(z < 3):.3f} within 3 se", file=sys.stderr, flush=True)
    return {"options": 512, "greek_mode": "Full", "ms": ms, "device_ms": fr[0]["device_ms"],
            "mc_fp32_vs_fourier_max_abs_z": float(z.max()), "mc_fp32_vs_fourier_frac_within_3se": float(np.mean(z < 3)),
            "mc_fp32_vs_fourier_rms_z": float(np.sqrt(np.mean(z * z))),
            "mc_fp32_vs_fourier_mean_z": float(np.mean(zs))}


def mlmc_leg(eng):
    """Multilevel MC of an up-and-out call under Heston (kappa 2, theta 0.04,
    sigma 0.3, rho -0.7, v0 0.04; SchemeNV), K 100, B 130, T 1: levels 0..6
    with 4 * 2^l fine steps, epsilon = 1e-4 x the price of a plain pilot run,
    Giles' sample sizes, FP32 paths, FP64 level sums, seed 42; each level a
    replayed HIP graph. 262,144 streams: one lane each, four wavefronts per
    SIMD (the level kernel is compiled for four), so a level launch fills the
    chip; the library default of 65,536 leaves it at one."""
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, mc_config, mlmc_config
    m = make_model(1, [Heston(2.0, 0.04, 0.3, -0.7, 0.04)])
    inst = InstrumentSpec(kind=capi.INSTR_BARRIER, strikes=[100.0], barrier=130.0, maturity=1.0,
                          barrier_dir=capi.BARRIER_UP, barrier_action=capi.BARRIER_OUT)
    pilot = eng.mc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE,
                   mc_config(1 << 20, 256, seed=1, precision=capi.FP32), S0, R, Q)
    eps = 1e-4 * float(pilot["price"][0])
    streams = 4 * 65536
    cfg = mlmc_config(epsilon=eps, L_min=6, L_max=6, N0=10000, base_steps=4, seed=42,
                      precision=capi.FP32, num_streams=streams)
    eng.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE,
             mlmc_config(epsilon=100 * eps, L_min=6, L_max=6, N0=1000, base_steps=4, seed=3,
                         precision=capi.FP32, num_streams=streams), S0, R, Q)  # warm-up
    t0 = time.perf_counter()
    res = eng.mlmc(m, [capi.SCHEME_NV], inst, capi.GREEK_NONE, cfg, S0, R, Q)
    wall = 1e3 * (time.perf_counter() - t0)
    print(f"# mlmc up-and-out call: {wall:.0f} ms, price {res['price']:.6f} +- "
          f"{res['price_std_err']:.2e}, {res['total_steps']:.3e} path-steps", file=sys.stderr,
          flush=True)
    return {"workload": "Heston NV up-and-out call K 100 B 130 T 1, levels 0..6, base 4 steps, "
                        "eps 1e-4 x pilot price, FP32 paths, FP64 sums, seed 42, 262,144 streams",
            "epsilon": eps, "price": res["price"], "price_std_err": res["price_std_err"],
            "ms": wall, "device_ms": res["device_ms"], "total_path_steps": res["total_steps"],
            "path_steps_per_s": res["total_steps"] / (wall / 1e3),
            "roofline": {"bound": "valu", "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "achieved": res["total_steps"] * FLOP_PER_PATH_STEP_MLMC / (res["device_ms"] / 1e3) / 1e12,
                         "frac": res["total_steps"] * FLOP_PER_PATH_STEP_MLMC / (res["device_ms"] / 1e3) / 1e12
                         / FP32_PEAK_TFLOPS,
                         "flop_per_path_step": FLOP_PER_PATH_STEP_MLMC},
            "paths_per_level": [int(x) for x in res["paths_per_level"]],
            "alpha": res["alpha_estimated"], "beta": res["beta_estimated"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--paths-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cv", action="store_true", help="skip the price-only and control-variate legs")
    ap.add_argument("--no-mlmc", action="store_true", help="skip the multilevel leg")
    ap.add_argument("--no-qmc", action="store_true", help="skip the quasi-Monte Carlo leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched Double-Heston leg")
    ap.add_argument("--no-fourier", action="store_true", help="skip the Gil-Pelaez leg")
    ap.add_argument("--no-ref-rows", action="store_true",
                    help="skip the strip timed under the reference's NCI semantics")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the strip kernel at its default launch shape (256-stream units, "
                         "strike loop unrolled by 4, grid 4x) instead of autotuning it")
    ap.add_argument("--shape", default=None,
                    help="BLOCK,UNROLL,GRID[,FOUR_WAVE]: fix the strip kernel's launch shape (no "
                         "autotuning), e.g. to profile the shape an autotuned run chose")
    ap.add_argument("--batched-by-scheme", action="store_true",
                    help="only time the batched Double-Heston sets split by their SchemeExact factor "
                         "count, print one JSON line and exit")
    ap.add_argument("--cv-net-save", default=None, metavar="PATH",
                    help="only train the three control-variate legs' networks, write them to PATH.<i> "
                         "and exit")
    ap.add_argument("--cv-net-load", default=None, metavar="PATH",
                    help="only price the three control-variate legs with the networks in PATH.<i> "
                         "(no training), print one JSON line and exit")
    ap.add_argument("--cv-jump-tables", action="store_true",
                    help="only build the control variate's terminal-jump tables of the Bates-Kou strip "
                         "(vanilla and smoothed digitals), print one JSON line and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: sharding the strip's streams "
                                   "across GPUs is designed but not launched; use --gpus 1"}))
        sys.exit(2)

    from myquantlib_amd import capi
    from myquantlib_amd.pricing import Engine
    eng = Engine(0)
    if args.batched_by_scheme:
        print(json.dumps({"batched_by_scheme": batched_by_scheme(eng)}))
        return
    if args.cv_net_save or args.cv_net_load:
        print(json.dumps({"cv_nets": cv_nets(eng, args.cv_net_save or args.cv_net_load, bool(args.cv_net_save))}))
        return
    if args.cv_jump_tables:
        print(json.dumps({"cv_jump_tables": cv_jump_tables(eng)}))
        return
    paths = 1 << args.paths_log2
    flags = 0 if (args.no_autotune or args.shape) else capi.FLAG_AUTOTUNE
    if args.shape:
        parts = [int(x) for x in args.shape.split(",")]
        blk, unr, gf = parts[:3]
        w4 = bool(parts[3]) if len(parts) > 3 else False
        for prec in (capi.FP32, capi.FP64):
            eng.set_launch_shape(prec, 1, 5, False, blk, unr, gf, w4)
    J32 = jobs(paths, capi.FP32, flags=flags)
    if flags:
        # tuning pass (untimed): trial launches of every shape at 1/8 of the
        # samples, then the call itself at the chosen shape
        t0 = time.perf_counter()
        eng.mc_batch(J32)
        print(f"# autotune fp32: {time.perf_counter() - t0:.3f} s; {eng.tuning_report().strip()}",
              file=sys.stderr, flush=True)
    for i in range(args.warmup):
        t0 = time.perf_counter()
        eng.mc_batch(J32)
        print(f"# warmup {i}: {time.perf_counter() - t0:.3f} s", file=sys.stderr, flush=True)
    dev_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = eng.mc_batch(J32)
        dev_ms.append(res[0]["device_ms"])
        print(f"# step {i}: {time.perf_counter() - t0:.3f} s elapsed, device {dev_ms[-1]:.1f} ms",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    total = path_steps(paths)
    value = total / (wall / args.steps)
    t32, worst32 = ms_to_target(res, ms_per_step)
    print(f"# fp32 strip: {ms_per_step:.1f} ms/step, device {np.mean(dev_ms):.1f} ms, "
          f"worst rel se {worst32:.2e}", file=sys.stderr, flush=True)

    fourier_leg = None if args.no_fourier else fourier_check(eng, res)
    ref_rows = None if args.no_ref_rows else reference_rows_leg(eng, paths, flags, fp64=not args.no_fp64)

    ms64 = None
    fp32_err = None
    if not args.no_fp64:
        J64 = jobs(paths, capi.FP64, flags=flags)
        eng.mc_batch(J64)
        t0 = time.perf_counter()
        r64 = eng.mc_batch(J64)
        w64 = 1e3 * (time.perf_counter() - t0)
        ms64, worst64 = ms_to_target(r64, w64)
        print(f"# fp64 strip: {w64:.1f} ms, worst rel se {worst64:.2e}", file=sys.stderr, flush=True)
        fp32_err = fp32_vs_fp64(res, r64)
        print(f"# fp32 paths vs fp64 paths on the same draws: {fp32_err}", file=sys.stderr, flush=True)

    # price-only strip (GreekMode None): plain FP32 / FP64 and the learned CV
    cv = {}
    po32 = po64 = None
    if not args.no_cv:
        Jn = jobs(paths, capi.FP32, mode=capi.GREEK_NONE, flags=flags)
        eng.mc_batch(Jn)  # warm-up (and launch shape of the price-only kernel)
        t0 = time.perf_counter()
        rn = eng.mc_batch(Jn)
        po32, _ = ms_to_target(rn, 1e3 * (time.perf_counter() - t0))
        if not args.no_fp64:
            Jn64 = jobs(paths, capi.FP64, mode=capi.GREEK_NONE, flags=flags)
            eng.mc_batch(Jn64)
            t0 = time.perf_counter()
            rn64 = eng.mc_batch(Jn64)
            po64, _ = ms_to_target(rn64, 1e3 * (time.perf_counter() - t0))
        cv_paths = min(paths, 1 << 22)
        cv["bates_strip"] = cv_leg(eng, cv_paths, bates_kou=False)
        cv["bates_kou_nv_strip"] = cv_leg(eng, cv_paths, bates_kou=True)
        cv["bates_kou_nv_smoothed_digitals"] = cv_leg(eng, cv_paths, bates_kou=True, payoff=capi.PAYOFF_CASH)
    qmc = None if (args.no_cv or args.no_qmc) else qmc_leg(eng, min(paths, 1 << 22), po32)

    mlmc = None if args.no_mlmc else mlmc_leg(eng)
    batched = None if args.no_batched else batched_leg(eng, flags=flags)

    kern_s = np.mean(dev_ms) / 1e3
    achieved = total * FLOP_PER_PATH_STEP / kern_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "bates_strip: Bates SchemeNCI (Feller violated) + Merton, 64 strikes x 8 "
                        f"maturities, 2^{args.paths_log2} paths per maturity on {STRIP_STREAMS} streams, "
                        "100 steps/yr, GreekMode::Full CRN, FP32 paths, FP64 sums, seed 42",
            "nci_semantics": "engine default: NCI rows to lambda(max(v0, 25 theta)) + 6 sqrt(lambda), upper-tail "
                             "table for u >= 0.999, first grid interval inverted exactly for rows N >= 2; the "
                             "reference's rows and fallback (VolSchemes.hpp:245-255,334-370) are timed under "
                             "reference_nci_rows",
            "paths_per_maturity": paths,
            "path_steps_per_step": total,
            "launch_shape": ("autotuned: " + eng.tuning_report().strip()) if flags else
                            (f"fixed {args.shape} (block,unroll,grid)" if args.shape else "default 256,4,4"),
        },
        # price-only runs of the strip's 512 options (GreekMode None); the CV
        # time includes its training; "full_greeks_*" are the Full-mode runs
        # FP32-path error against FP64 paths on the same draws, per field
        "fp32_vs_fp64_same_draws_max_rel": fp32_err,
        "ms_to_1e-4_rel_se": {
            "plain_fp32": po32, "plain_fp64": po64,
            "learned_cv_fp32": cv["bates_strip"]["learned_cv_fp32_ms_to_1e-4"] if cv else None,
            "learned_cv_speedup_vs_plain_fp64": cv["bates_strip"]["speedup_vs_plain_fp64"] if cv else None,
            "full_greeks_fp32": t32, "full_greeks_fp64": ms64},
        "learned_cv": cv or None,
        "qmc": qmc,
        "mlmc": mlmc,
        "batched_double_heston": batched,
        "fourier": fourier_leg,
        "reference_nci_rows": ref_rows,
        "sample_steps_per_s": value / VARIANTS,
        "executed_variant_fraction": EXECUTED_VARIANTS / VARIANTS,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH.get(args.paths_log2),
                     "kernel_ms": float(np.mean(dev_ms)),
                     "flop_per_path_step": FLOP_PER_PATH_STEP},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
