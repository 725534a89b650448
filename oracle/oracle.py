"""ORACLE (test infrastructure only): ctypes front-end of the CPU FP64
restatement in oracle/src (lib/libmyql_oracle.so). Only tests/, the smoke
check in __graft_entry__.py and bench.py's cpu_baseline leg import this
module, and only as the checker or the CPU baseline.

The entry points mirror the product's C-ABI so a parity test calls both with
the very same structs (myquantlib_amd.capi).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from myquantlib_amd import capi
from myquantlib_amd.pricing import (ResultBuffers, MLMCBuffers, schemes_array, InstrumentSpec,
                                    EngineError)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libmyql_oracle.so")
LIB_FAST = os.path.join(HERE, "lib", "libmyql_oracle_fast.so")
LIB_OPS = os.path.join(HERE, "lib", "libmyql_oracle_ops.so")


def build(quiet: bool = True) -> None:
    """Compile the oracle libraries with the committed Makefile."""
    out = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


class Oracle:
    def __init__(self, fast: bool = False, ops: bool = False):
        """fast: the reference-flags build (CPU baseline); ops: the op-counting
        build (flops per path-step; run it with nthreads=1)."""
        path = LIB_OPS if ops else (LIB_FAST if fast else LIB)
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        u64p = C.POINTER(C.c_uint64)
        dp = C.POINTER(C.c_double)
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_op_count.restype = C.c_uint64
        lib.oracle_op_count.argtypes = [C.c_int]
        lib.oracle_seed_state.argtypes = [C.c_uint64, u64p]
        lib.oracle_next.argtypes = [u64p, C.c_int64, u64p]
        lib.oracle_char_poly.argtypes = [u64p]
        lib.oracle_jump_poly.argtypes = [C.c_int, u64p]
        lib.oracle_apply_poly.argtypes = [u64p, u64p]
        lib.oracle_x_pow_mod.argtypes = [C.c_uint64, u64p]
        lib.oracle_x_pow2_mod.argtypes = [C.c_int, u64p]
        lib.oracle_stream_state.argtypes = [C.c_uint64, C.c_int, C.c_uint64, u64p]
        for f in ("oracle_normcdfinv",):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.c_double]
        for f in ("oracle_poissinvcdf", "oracle_chi2_quantile", "oracle_gamma_p", "oracle_gamma_q"):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.c_double, C.c_double]
        lib.oracle_uniforms.argtypes = [u64p, C.c_int64, dp]
        lib.oracle_poisson_samples.argtypes = [C.c_double, C.c_int64, C.c_uint64,
                                               C.POINTER(C.c_int32)]
        lib.oracle_cir_terminal.restype = C.c_int
        lib.oracle_cir_terminal.argtypes = [C.c_int32, C.POINTER(capi.HestonParams), C.c_double,
                                            C.c_int64, C.c_int64, C.c_uint64, dp, C.c_int]
        lib.oracle_path_payoffs.restype = C.c_int
        lib.oracle_path_payoffs.argtypes = [C.POINTER(capi.Instrument), C.c_int32, C.c_double, C.c_double, dp,
                                            C.c_int64, C.c_double, dp, dp, dp]
        lib.oracle_nci_lookup.restype = C.c_int
        lib.oracle_nci_lookup.argtypes = [C.POINTER(capi.HestonParams), C.c_double, C.c_int, C.c_int32, dp,
                                          C.c_int64, dp]
        lib.oracle_mc_calculate.restype = C.c_int
        lib.oracle_mc_calculate.argtypes = [C.POINTER(capi.Model), C.POINTER(C.c_int32),
                                            C.POINTER(capi.Instrument), C.c_int32,
                                            C.POINTER(capi.MCConfig), C.c_double, C.c_double,
                                            C.c_double, C.POINTER(capi.MCResult), C.c_int]
        for fn in (lib.oracle_debug_plain_path, lib.oracle_debug_coupled_path):
            fn.restype = C.c_int
        lib.oracle_debug_plain_path.argtypes = [C.POINTER(capi.Model), C.POINTER(C.c_int32),
                                                C.POINTER(capi.Instrument), C.c_uint64, C.c_uint64, C.c_double,
                                                C.c_double, C.c_double, dp, dp, dp, C.c_int64,
                                                C.POINTER(C.c_int64)]
        lib.oracle_debug_coupled_path.argtypes = [C.POINTER(capi.Model), C.POINTER(C.c_int32),
                                                  C.POINTER(capi.Instrument), C.c_uint64, C.c_uint64, C.c_double,
                                                  C.c_double, C.c_double, dp, dp, dp, dp, C.c_int64,
                                                  C.POINTER(C.c_int64)]
        lib.oracle_mlmc_calculate.restype = C.c_int
        lib.oracle_mlmc_calculate.argtypes = [C.POINTER(capi.Model), C.POINTER(C.c_int32),
                                              C.POINTER(capi.Instrument), C.c_int32,
                                              C.POINTER(capi.MLMCConfig), C.c_double, C.c_double,
                                              C.c_double, C.POINTER(capi.MLMCResult), C.c_int]
        lib.oracle_mlmc_level_weights.restype = C.c_int
        lib.oracle_mlmc_level_weights.argtypes = [C.c_int64, C.c_double, C.c_int32, dp]
        self.lib = lib

    def mlmc_level_weights(self, L: int, alpha: float, ml2r: bool = True) -> np.ndarray:
        """Weights W_0..W_L of the level means (ML2R: from the Vandermonde
        system of the Richardson weights, solved by elimination)."""
        W = np.zeros(L + 1)
        rc = self.lib.oracle_mlmc_level_weights(L, alpha, int(ml2r), W.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise EngineError(rc, "bad arguments")
        return W

    def op_count(self, reset: bool = False) -> int:
        """Floating-point ops counted since the last reset (ops build only)."""
        n = int(self.lib.oracle_op_count(int(reset)))
        if n == 2 ** 64 - 1:
            raise RuntimeError("op counting needs Oracle(ops=True)")
        return n

    # ---- RNG -------------------------------------------------------------
    @staticmethod
    def _u64(n):
        a = np.zeros(n, dtype=np.uint64)
        return a, a.ctypes.data_as(C.POINTER(C.c_uint64))

    def seed_state(self, seed: int) -> np.ndarray:
        a, p = self._u64(4)
        self.lib.oracle_seed_state(seed, p)
        return a

    def next(self, state: np.ndarray, n: int) -> np.ndarray:
        st = np.ascontiguousarray(state, dtype=np.uint64).copy()
        out, po = self._u64(n)
        self.lib.oracle_next(st.ctypes.data_as(C.POINTER(C.c_uint64)), n, po)
        return out

    def char_poly(self) -> np.ndarray:
        a, p = self._u64(8)
        self.lib.oracle_char_poly(p)
        return a

    def jump_poly(self, k: int) -> np.ndarray:
        a, p = self._u64(4)
        self.lib.oracle_jump_poly(k, p)
        return a

    def x_pow_mod(self, n: int) -> np.ndarray:
        a, p = self._u64(4)
        self.lib.oracle_x_pow_mod(n, p)
        return a

    def x_pow2_mod(self, e: int) -> np.ndarray:
        a, p = self._u64(4)
        self.lib.oracle_x_pow2_mod(e, p)
        return a

    def apply_poly(self, state: np.ndarray, poly: np.ndarray) -> np.ndarray:
        st = np.ascontiguousarray(state, dtype=np.uint64).copy()
        pl = np.ascontiguousarray(poly, dtype=np.uint64)
        self.lib.oracle_apply_poly(st.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   pl.ctypes.data_as(C.POINTER(C.c_uint64)))
        return st

    def stream_state(self, seed: int, layout: int, s: int) -> np.ndarray:
        a, p = self._u64(4)
        self.lib.oracle_stream_state(seed, layout, s, p)
        return a

    def uniforms(self, state: np.ndarray, n: int) -> np.ndarray:
        st = np.ascontiguousarray(state, dtype=np.uint64).copy()
        out = np.zeros(n)
        self.lib.oracle_uniforms(st.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                                 out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    # ---- numerics --------------------------------------------------------
    def normcdfinv(self, p: float) -> float:
        return self.lib.oracle_normcdfinv(p)

    def poissinvcdf(self, lam: float, u: float) -> float:
        return self.lib.oracle_poissinvcdf(lam, u)

    def chi2_quantile(self, dof: float, p: float) -> float:
        return self.lib.oracle_chi2_quantile(dof, p)

    def gamma_p(self, a: float, x: float) -> float:
        return self.lib.oracle_gamma_p(a, x)

    def gamma_q(self, a: float, x: float) -> float:
        return self.lib.oracle_gamma_q(a, x)

    def poisson_samples(self, mean: float, n: int, seed: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.int32)
        self.lib.oracle_poisson_samples(mean, n, seed, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def path_payoffs(self, instr, mode: int, S0: float, fd_bump: float, logS, dt: float):
        """Trackers and payoffs of `instr` (an InstrumentSpec) on the log-price
        path logS[0..n]: (base, up, dn) arrays over the strikes."""
        st = instr.struct()
        logS = np.ascontiguousarray(logS, dtype=np.float64)
        n = st.n_strikes
        base, up, dn = np.zeros(n), np.zeros(n), np.zeros(n)
        pd = C.POINTER(C.c_double)
        rc = self.lib.oracle_path_payoffs(C.byref(st), mode, S0, fd_bump, logS.ctypes.data_as(pd), logS.size - 1, dt,
                                          base.ctypes.data_as(pd), up.ctypes.data_as(pd), dn.ctypes.data_as(pd))
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        return base, up, dn

    def nci_lookup(self, h, dt: float, N: int, u, reference_rows: bool = False) -> np.ndarray:
        """The NCI table's chi2_{d+2N} quantile (before scaling) at uniforms u."""
        p = capi.HestonParams(h.kappa, h.theta, h.sigma, h.rho, h.v0)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(u.size)
        rc = self.lib.oracle_nci_lookup(C.byref(p), dt, int(reference_rows), N,
                                        u.ctypes.data_as(C.POINTER(C.c_double)), u.size,
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        return out

    def cir_terminal(self, scheme: int, h, dt: float, n_steps: int, n: int, seed: int = 1,
                     nthreads: int = 0) -> np.ndarray:
        p = capi.HestonParams(h.kappa, h.theta, h.sigma, h.rho, h.v0)
        out = np.zeros(n)
        rc = self.lib.oracle_cir_terminal(scheme, C.byref(p), dt, n_steps, n, seed,
                                          out.ctypes.data_as(C.POINTER(C.c_double)), nthreads)
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        return out

    # ---- pricers ---------------------------------------------------------
    def mc(self, model, schemes, instr: InstrumentSpec, mode: int, cfg, S0: float, r: float,
           q: float, nthreads: int = 0) -> dict:
        s = instr.struct()
        buf = ResultBuffers(s.n_strikes)
        rc = self.lib.oracle_mc_calculate(C.byref(model), schemes_array(schemes), C.byref(s), mode,
                                          C.byref(cfg), S0, r, q, C.byref(buf.struct), nthreads)
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        return buf.as_dict()

    def mlmc(self, model, schemes, instr: InstrumentSpec, mode: int, cfg, S0: float, r: float,
             q: float, nthreads: int = 0) -> dict:
        s = instr.struct()
        buf = MLMCBuffers(int(cfg.L_max))
        rc = self.lib.oracle_mlmc_calculate(C.byref(model), schemes_array(schemes), C.byref(s),
                                            mode, C.byref(cfg), S0, r, q, C.byref(buf.struct),
                                            nthreads)
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        return buf.as_dict()

    PLAIN_REC, COUPLED_REC, FIN = 5, 16, 10

    def debug_plain_path(self, model, schemes, instr: InstrumentSpec, steps: int, seed: int, S0: float, r: float,
                         q: float):
        """One plain-stepper path from seed_state(seed): per-step record
        (first-draw index, log S, V0, V1, step jump; plus the European's
        terminal jump row), the final tracker state and every uniform drawn."""
        rec = np.zeros((steps + 1) * self.PLAIN_REC)
        fin = np.zeros(self.FIN)
        draws = np.zeros(64 * steps + 1024)
        n = C.c_int64(0)
        s = instr.struct()
        rc = self.lib.oracle_debug_plain_path(C.byref(model), schemes_array(schemes), C.byref(s), steps, seed, S0,
                                              r, q, _dp(rec), _dp(fin), _dp(draws), draws.size, C.byref(n))
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        assert n.value <= draws.size
        return rec.reshape(steps + 1, self.PLAIN_REC), fin, draws[:n.value]

    def debug_coupled_path(self, model, schemes, instr: InstrumentSpec, coarse_steps: int, seed: int, S0: float,
                           r: float, q: float):
        """One MLMC coupled path (fine step T / (2 coarse_steps)) from
        seed_state(seed): per-coarse-step record (first-draw index, G1, G2,
        Gc, a1, a2, j1, j2, xf, vf0, vf1, xc, vc0, vc1, mid0, mid1; plus the
        European's shared jump row), the final fine and coarse tracker
        states and every uniform drawn."""
        nc = coarse_steps
        rec = np.zeros((nc + 1) * self.COUPLED_REC)
        ff, fc = np.zeros(self.FIN), np.zeros(self.FIN)
        draws = np.zeros(128 * nc + 1024)
        n = C.c_int64(0)
        s = instr.struct()
        rc = self.lib.oracle_debug_coupled_path(C.byref(model), schemes_array(schemes), C.byref(s), nc, seed, S0,
                                                r, q, _dp(rec), _dp(ff), _dp(fc), _dp(draws), draws.size,
                                                C.byref(n))
        if rc != 0:
            raise EngineError(rc, self.lib.oracle_last_error().decode())
        assert n.value <= draws.size
        return rec.reshape(nc + 1, self.COUPLED_REC), ff, fc, draws[:n.value]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))
