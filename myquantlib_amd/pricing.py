"""Python host-side mirror of the engine's C-ABI: builders for the model,
instrument and config structs, result marshalling into numpy, and `Engine`,
a thin owner of one `myql_ctx` on a HIP device.

Names follow the reference library (HestonParams, MertonParams, KouParams,
MonteCarloConfig, MLMCConfig, GreekMode, EuropeanOption, ...): reference
include/myql/models/asvj/data/ModelParams.hpp:7-47,
include/myql/pricers/montecarlo/MonteCarloPricer.hpp:16-26,
include/myql/pricers/mlmc/MultiLevelMonteCarloPricer.hpp:16-30.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import capi
from .capi import (Model, Instrument, MCConfig, MCResult, MLMCConfig, MLMCResult,
                   HestonParams)

RESULT_FIELDS = ["price", "delta", "gamma", "theta", "rho", "charm"]
PAIR_FIELDS = ["vega", "vanna"]


@dataclass
class Heston:
    kappa: float
    theta: float
    sigma: float
    rho: float
    v0: float


def make_model(n_factors: int, heston: Sequence[Heston] = (), vol: float = 0.0,
               jump_kind: int = capi.JUMP_NONE, jump: Sequence[float] = (0, 0, 0, 0)) -> Model:
    m = Model()
    m.n_factors = n_factors
    m.jump_kind = jump_kind
    m.vol = vol
    for i, h in enumerate(heston):
        m.heston[i] = HestonParams(h.kappa, h.theta, h.sigma, h.rho, h.v0)
    for i, x in enumerate(list(jump) + [0.0] * (4 - len(jump))):
        m.jump[i] = x
    return m


def merton(lam: float, mu: float, delta: float):
    return capi.JUMP_MERTON, (lam, mu, delta, 0.0)


def kou(lam: float, p_up: float, eta1: float, eta2: float):
    return capi.JUMP_KOU, (lam, p_up, eta1, eta2)


@dataclass
class InstrumentSpec:
    """Owns the strike buffer the C struct points to."""
    kind: int = capi.INSTR_EUROPEAN
    payoff: int = capi.PAYOFF_VANILLA
    option_type: int = capi.CALL
    strikes: Sequence[float] = (100.0,)
    maturity: float = 1.0
    is_strip: Optional[bool] = None
    fixed_strike: bool = True
    barrier_dir: int = capi.BARRIER_UP
    barrier_action: int = capi.BARRIER_OUT
    rebate_timing: int = capi.REBATE_NONE
    barrier: float = 0.0
    rebate: float = 0.0
    rebate_rate: float = 0.0
    payoff_eps: float = -1.0
    past_avg_log: float = 0.0
    past_avg_S: float = 0.0
    t_elapsed: float = 0.0
    hist_min_S: float = 0.0
    hist_max_S: float = 0.0
    _k: np.ndarray = field(default=None, repr=False)
    _s: Instrument = field(default=None, repr=False)

    def struct(self) -> Instrument:
        """A C struct for this instrument. Each struct keeps its own strike
        array alive, so one spec can back several jobs of a batch."""
        k = np.ascontiguousarray(np.asarray(self.strikes, dtype=np.float64).reshape(-1))
        self._k = k
        s = Instrument()
        s._strikes_ref = k
        s.kind, s.payoff, s.option_type = self.kind, self.payoff, self.option_type
        s.fixed_strike = int(self.fixed_strike)
        s.barrier_dir, s.barrier_action = self.barrier_dir, self.barrier_action
        s.rebate_timing = self.rebate_timing
        strip = self.is_strip if self.is_strip is not None else len(self._k) > 1
        s.is_strip = int(strip)
        s.maturity, s.barrier, s.rebate = self.maturity, self.barrier, self.rebate
        s.rebate_rate, s.payoff_eps = self.rebate_rate, self.payoff_eps
        s.past_avg_log, s.past_avg_S, s.t_elapsed = self.past_avg_log, self.past_avg_S, self.t_elapsed
        s.hist_min_S, s.hist_max_S = self.hist_min_S, self.hist_max_S
        s.strikes = k.ctypes.data_as(C.POINTER(C.c_double))
        s.n_strikes = len(k)
        self._s = s
        return s


def mc_config(num_paths: int, time_steps: int, seed: int = 42, fd_bump: float = 1e-4,
              vol_bump: float = 1e-4, t_bump: float = 1e-4, r_bump: float = 1e-4,
              precision: int = capi.FP64, stream_layout: int = capi.STREAMS_JUMP,
              num_streams: int = 0, flags: int = 0, qmc_replicates: int = 0) -> MCConfig:
    return MCConfig(num_paths, time_steps, seed, fd_bump, vol_bump, t_bump, r_bump, precision,
                    stream_layout, num_streams, flags, qmc_replicates)


def cv_config(train_paths: int = 1 << 20, train_seed: int = 7, steps: int = 2000, batch: int = 8192,
              lr: float = 1e-3, points: int = 8, init_seed: int = 1) -> "capi.CVConfig":
    return capi.CVConfig(train_paths, train_seed, steps, batch, lr, points, init_seed)


def mlmc_config(epsilon: float = 0.005, L_min: int = 2, L_max: int = 15, N0: int = 10000,
                base_steps: int = 1, alpha: float = 2.0, beta: float = 1.0,
                adaptive_rates: bool = True, seed: int = 42, fd_bump: float = 1e-4,
                precision: int = capi.FP64, stream_layout: int = capi.STREAMS_JUMP,
                num_streams: int = 0, flags: int = 0, estimator: int = capi.MLMC_GILES) -> MLMCConfig:
    return MLMCConfig(epsilon, L_min, L_max, N0, base_steps, alpha, beta, int(adaptive_rates),
                      seed, fd_bump, precision, stream_layout, num_streams, flags, estimator)


class ResultBuffers:
    """numpy arrays that an MCResult points into: one NaN-filled block, one
    row per result array (two for the per-factor fields)."""

    ROWS = 2 * len(RESULT_FIELDS) + 4 * len(PAIR_FIELDS)

    def __init__(self, n: int):
        self.n = n
        self.arr = {}
        block = np.full((self.ROWS, n), np.nan)
        self._block = block
        base, stride = block.ctypes.data, block.strides[0]
        r = MCResult()
        row = 0
        for f in RESULT_FIELDS:
            for suffix in ("", "_std_err"):
                self.arr[f + suffix] = block[row]
                setattr(r, f + suffix, base + row * stride)
                row += 1
        for f in PAIR_FIELDS:
            for suffix in ("", "_std_err"):
                self.arr[f + suffix] = block[row:row + 2]
                arrp = getattr(r, f + suffix)
                arrp[0] = base + row * stride
                arrp[1] = base + (row + 1) * stride
                row += 2
        self.struct = r

    def as_dict(self) -> dict:
        d = {k: v.copy() for k, v in self.arr.items()}
        d["total_steps"] = int(self.struct.total_steps)
        d["device_ms"] = float(self.struct.device_ms)
        return d


class MLMCBuffers:
    def __init__(self, L_max: int):
        self.ppl = np.zeros(L_max + 1, dtype=np.uint64)
        self.mean = np.full(L_max + 1, np.nan)
        self.var = np.full(L_max + 1, np.nan)
        self.weights = np.full(L_max + 1, np.nan)
        r = MLMCResult()
        r.paths_per_level = self.ppl.ctypes.data_as(C.POINTER(C.c_uint64))
        r.mean_diffs = self.mean.ctypes.data_as(C.POINTER(C.c_double))
        r.var_diffs = self.var.ctypes.data_as(C.POINTER(C.c_double))
        r.level_weights = self.weights.ctypes.data_as(C.POINTER(C.c_double))
        self.struct = r

    def as_dict(self) -> dict:
        r = self.struct
        n = int(r.n_levels)
        return dict(price=r.price, price_std_err=r.price_std_err, total_steps=int(r.total_steps),
                    n_levels=n, paths_per_level=self.ppl[:n].copy(),
                    mean_diffs=self.mean[:n].copy(), var_diffs=self.var[:n].copy(),
                    level_weights=self.weights[:n].copy(),
                    delta=r.delta, delta_std_err=r.delta_std_err, gamma=r.gamma, gamma_std_err=r.gamma_std_err,
                    alpha_estimated=r.alpha_estimated, beta_estimated=r.beta_estimated,
                    device_ms=r.device_ms)


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def schemes_array(schemes: Sequence[int]):
    s = list(schemes) + [capi.SCHEME_NULL] * (2 - len(schemes))
    return (C.c_int32 * 2)(*s)


class Engine:
    """Owner of one device context of libmyql_hip.so."""

    def __init__(self, device: int = 0, lib_path: str = capi.LIB_PATH):
        self.lib = capi.load_library(lib_path)
        self.ctx = self.lib.myql_create(device)
        if not self.ctx:
            raise EngineError(capi.E_HIP, "myql_create failed (no usable HIP device?)")

    def close(self):
        if self.ctx:
            self.lib.myql_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != capi.OK:
            msg = self.lib.myql_last_error(self.ctx).decode()
            raise EngineError(rc, msg)

    def set_launch_shape(self, precision: int, n_factors: int, lockstep_width: int, path_dependent: bool,
                         block: int, unroll: int, grid_factor: int = 4, four_wave: bool = False):
        """Fix one path kernel's launch shape (myql_debug_set_launch_shape)."""
        self._check(self.lib.myql_debug_set_launch_shape(self.ctx, precision, n_factors, lockstep_width,
                                                         int(path_dependent), block, unroll, grid_factor,
                                                         int(four_wave)))

    def tuning_report(self) -> str:
        """Launch shapes chosen under FLAG_AUTOTUNE on this context (one line
        per kernel, with the trial times); empty before any tuned call."""
        buf = C.create_string_buffer(8192)
        self._check(self.lib.myql_tuning_report(self.ctx, buf, len(buf)))
        return buf.value.decode()

    def mc(self, model: Model, schemes, instr: InstrumentSpec, mode: int, cfg: MCConfig,
           S0: float, r: float, q: float) -> dict:
        s = instr.struct()
        buf = ResultBuffers(s.n_strikes)
        rc = self.lib.myql_mc_calculate(self.ctx, C.byref(model), schemes_array(schemes),
                                        C.byref(s), mode, C.byref(cfg), S0, r, q,
                                        C.byref(buf.struct))
        self._check(rc)
        return buf.as_dict()

    def mc_noalloc(self, model: Model, schemes, instr: InstrumentSpec, mode: int, cfg: MCConfig,
                   S0: float, r: float, q: float) -> dict:
        """mc() through myql_mc_calculate_noalloc: the call's device memory
        comes from the library's per-GPU scratch array, never an allocation."""
        s = instr.struct()
        buf = ResultBuffers(s.n_strikes)
        rc = self.lib.myql_mc_calculate_noalloc(self.ctx, C.byref(model), schemes_array(schemes),
                                                C.byref(s), mode, C.byref(cfg), S0, r, q,
                                                C.byref(buf.struct))
        self._check(rc)
        return buf.as_dict()

    @staticmethod
    def _job_array(jobs: list):
        keep, arr, bufs = [], (capi.MCJob * len(jobs))(), []
        structs = {}  # one C struct per instrument spec object of the batch
        for i, (model, schemes, instr, mode, cfg, S0, r, q) in enumerate(jobs):
            s = structs.get(id(instr))
            if s is None:
                s = structs[id(instr)] = instr.struct()
            b = ResultBuffers(s.n_strikes)
            keep.append((model, s, cfg, instr))
            bufs.append(b)
            j = arr[i]
            j.model = C.pointer(model)
            j.schemes = schemes_array(schemes)
            j.greek_mode = mode
            j.instrument = C.pointer(s)
            j.config = C.pointer(cfg)
            j.S0, j.r, j.q = S0, r, q
            j.result = C.pointer(b.struct)
        return arr, bufs, keep

    def mc_batch(self, jobs: list) -> list:
        """jobs: list of (model, schemes, instr_spec, mode, cfg, S0, r, q)."""
        arr, bufs, _keep = self._job_array(jobs)
        rc = self.lib.myql_mc_calculate_batch(self.ctx, arr, len(jobs))
        self._check(rc)
        return [b.as_dict() for b in bufs]

    # ---- Gil-Pelaez Fourier pricer -----------------------------------------
    def fourier_batch(self, jobs: list, start_bound_guess: float = 20.0, tolerance: float = 1e-8) -> list:
        """jobs: list of (model, instr_spec, mode, S0, r, q) for European
        options; one launch for all strikes of all jobs. Result dicts carry
        price and the Greeks of the mode (std errors stay NaN)."""
        cfg = capi.FourierConfig(tolerance, start_bound_guess)
        arr = (capi.FourierJob * len(jobs))()
        bufs, keep = [], [cfg]
        for i, (model, instr, mode, S0, r, q) in enumerate(jobs):
            st = instr.struct()
            buf = ResultBuffers(len(instr.strikes))
            keep += [model, st]
            bufs.append(buf)
            arr[i].model = C.pointer(model)
            arr[i].instrument = C.pointer(st)
            arr[i].greek_mode = mode
            arr[i].config = C.pointer(cfg)
            arr[i].S0, arr[i].r, arr[i].q = S0, r, q
            arr[i].result = C.pointer(buf.struct)
        self._check(self.lib.myql_fourier_calculate_batch(self.ctx, arr, len(jobs)))
        return [b.as_dict() for b in bufs]

    def fourier(self, model, instr, mode: int, S0: float, r: float, q: float, **kw) -> dict:
        return self.fourier_batch([(model, instr, mode, S0, r, q)], **kw)[0]

    # ---- learned control variate -----------------------------------------
    def cv_train(self, jobs: list, cfg: "capi.CVConfig") -> "capi.CVNet":
        """Train the control-variate network for the option sets of `jobs`
        (same tuples as mc_batch; their path counts and seeds are unused)."""
        arr, _bufs, _keep = self._job_array(jobs)
        net = capi.CVNet()
        self._check(self.lib.myql_cv_train(self.ctx, arr, len(jobs), C.byref(cfg), C.byref(net)))
        return net

    def mc_batch_cv(self, jobs: list, cfg: "capi.CVConfig", net: "capi.CVNet") -> list:
        """Price `jobs` with the control variate; each result dict also holds
        plain_price, plain_std_err, var_ratio and used_cv per strike."""
        arr, bufs, _keep = self._job_array(jobs)
        extra, crs = [], (capi.CVResult * len(jobs))()
        pd = C.POINTER(C.c_double)
        for i, b in enumerate(bufs):
            e = {"plain_price": np.full(b.n, np.nan), "plain_std_err": np.full(b.n, np.nan),
                 "var_ratio": np.full(b.n, np.nan), "used_cv": np.zeros(b.n, dtype=np.int32)}
            extra.append(e)
            crs[i].plain_price = e["plain_price"].ctypes.data_as(pd)
            crs[i].plain_std_err = e["plain_std_err"].ctypes.data_as(pd)
            crs[i].var_ratio = e["var_ratio"].ctypes.data_as(pd)
            crs[i].used_cv = e["used_cv"].ctypes.data_as(C.POINTER(C.c_int32))
        self._check(self.lib.myql_mc_calculate_batch_cv(self.ctx, arr, len(jobs), C.byref(cfg),
                                                        C.byref(net), crs))
        out = []
        for b, e in zip(bufs, extra):
            d = b.as_dict()
            d.update({k: v.copy() for k, v in e.items()})
            out.append(d)
        return out

    def cv_input_grad(self, params: np.ndarray, feats: np.ndarray) -> np.ndarray:
        """Device (FP16 MFMA) tangents of the 64 outputs along features 1..3
        for rows of 5 features: array (n, 64, 3)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 5)
        out = np.zeros((f.shape[0], 64, 3), dtype=np.float32)
        pf = C.POINTER(C.c_float)
        self._check(self.lib.myql_debug_cv_input_grad(self.ctx, p.ctypes.data_as(pf), f.ctypes.data_as(pf),
                                                      f.shape[0], out.ctypes.data_as(pf)))
        return out

    def cv_param_grad(self, params: np.ndarray, feats: np.ndarray, targets: np.ndarray):
        """FP32 training gradient of the mean over rows and the 64 outputs of
        (net - target)^2: rows of 5 features, targets (n, 64)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 64)
        g = np.zeros(capi.CV_NPARAM, dtype=np.float32)
        loss = C.c_double(0.0)
        pf = C.POINTER(C.c_float)
        self._check(self.lib.myql_debug_cv_param_grad(self.ctx, p.ctypes.data_as(pf), f.ctypes.data_as(pf),
                                                      t.ctypes.data_as(pf), f.shape[0],
                                                      g.ctypes.data_as(pf), C.byref(loss)))
        return g, loss.value

    def mlmc(self, model: Model, schemes, instr: InstrumentSpec, mode: int, cfg: MLMCConfig,
             S0: float, r: float, q: float) -> dict:
        s = instr.struct()
        buf = MLMCBuffers(int(cfg.L_max))
        rc = self.lib.myql_mlmc_calculate(self.ctx, C.byref(model), schemes_array(schemes),
                                          C.byref(s), mode, C.byref(cfg), S0, r, q,
                                          C.byref(buf.struct))
        self._check(rc)
        return buf.as_dict()

    def stream_draws(self, seed: int, layout: int, n_streams: int, n_per: int) -> np.ndarray:
        out = np.zeros(n_streams * n_per, dtype=np.uint64)
        rc = self.lib.myql_debug_stream_draws(self.ctx, seed, layout, n_streams, n_per,
                                              out.ctypes.data_as(C.POINTER(C.c_uint64)))
        self._check(rc)
        return out.reshape(n_streams, n_per)

    def normcdfinv(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float64)
        o64 = np.zeros_like(p)
        o32 = np.zeros(p.shape, dtype=np.float32)
        pd = C.POINTER(C.c_double)
        rc = self.lib.myql_debug_normcdfinv(self.ctx, p.ctypes.data_as(pd), p.size,
                                            o64.ctypes.data_as(pd),
                                            o32.ctypes.data_as(C.POINTER(C.c_float)))
        self._check(rc)
        return o64, o32

    def poissinv(self, lam: np.ndarray, u: np.ndarray) -> np.ndarray:
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(lam)
        pd = C.POINTER(C.c_double)
        rc = self.lib.myql_debug_poissinv(self.ctx, lam.ctypes.data_as(pd), u.ctypes.data_as(pd),
                                          lam.size, out.ctypes.data_as(pd))
        self._check(rc)
        return out

    def poisson_counts(self, lam: np.ndarray, u: np.ndarray) -> np.ndarray:
        """The path kernels' Poisson counts: lam is (n, slots) with slots 1 or
        5 sharing the uniforms u (n,); returns int32 (n, slots)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        if lam.ndim == 1:
            lam = lam[:, None]
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(lam.shape, dtype=np.int32)
        pd = C.POINTER(C.c_double)
        rc = self.lib.myql_debug_poisson_counts(self.ctx, lam.ctypes.data_as(pd), lam.shape[1],
                                                u.ctypes.data_as(pd), u.size,
                                                out.ctypes.data_as(C.POINTER(C.c_int32)))
        self._check(rc)
        return out

    def qmc_points(self, seed: int, replicate: int, dims: int, n_points: int, n_streams: int) -> np.ndarray:
        """The randomized QMC coordinates the path kernels draw, as uint32
        (n_points, dims)."""
        out = np.zeros((n_points, dims), dtype=np.uint32)
        rc = self.lib.myql_debug_qmc_points(self.ctx, seed, replicate, dims, n_points, n_streams,
                                            out.ctypes.data_as(C.POINTER(C.c_uint32)))
        self._check(rc)
        return out

    def path_payoffs(self, instr: "InstrumentSpec", mode: int, S0: float, fd_bump: float, logS) -> tuple:
        """Device trackers and payoffs on log-price paths logS (n_paths x
        (n_steps + 1)): (base, up, dn), each n_paths x n_strikes."""
        logS = np.ascontiguousarray(np.atleast_2d(logS), dtype=np.float64)
        st = instr.struct()
        n_paths, n_steps = logS.shape[0], logS.shape[1] - 1
        out = [np.zeros((n_paths, st.n_strikes)) for _ in range(3)]
        pd = C.POINTER(C.c_double)
        rc = self.lib.myql_debug_path_payoffs(self.ctx, C.byref(st), mode, S0, fd_bump, logS.ctypes.data_as(pd),
                                              n_paths, n_steps, *[o.ctypes.data_as(pd) for o in out])
        self._check(rc)
        return tuple(out)

    def cv_jump_expectation(self, model: "Model", instr: "InstrumentSpec", mode: int, S0: float, spots) -> tuple:
        """The control variate's terminal-jump expectation E[P_k | S_T- = S]
        at the given spots (myql_debug_cv_jump_expectation): (tab, dir, ddir),
        each n_spots x n_strikes, or None where the estimator goes without it."""
        spots = np.ascontiguousarray(spots, dtype=np.float64)
        st = instr.struct()
        out = [np.zeros((spots.size, st.n_strikes)) for _ in range(3)]
        on = C.c_int32(0)
        pd = C.POINTER(C.c_double)
        self._check(self.lib.myql_debug_cv_jump_expectation(self.ctx, C.byref(model), C.byref(st), mode, S0,
                                                            spots.ctypes.data_as(pd), spots.size,
                                                            *[o.ctypes.data_as(pd) for o in out], C.byref(on)))
        return tuple(out) if on.value else None

    def chi2_quantile(self, dof: np.ndarray, p: np.ndarray) -> np.ndarray:
        dof = np.ascontiguousarray(dof, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.zeros_like(dof)
        pd = C.POINTER(C.c_double)
        rc = self.lib.myql_debug_chi2_quantile(self.ctx, dof.ctypes.data_as(pd),
                                               p.ctypes.data_as(pd), dof.size,
                                               out.ctypes.data_as(pd))
        self._check(rc)
        return out
