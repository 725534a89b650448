// Parameter blocks shared by the host orchestration (engine.hip) and the path
// kernels. Everything a kernel needs for one pricing call ("job") is packed
// into Job<R>, uploaded once per call and read through scalar loads.
#pragma once
#include <stdint.h>

namespace myql {

// Greek variants of one sample, in the order the reference replays them
// (MonteCarloPricer.hpp:306-409): base, v0 of factor 0 up/down, factor 1
// up/down, maturity up/down, rate up/down.
enum Variant { V_BASE = 0, V_V0UP, V_V0DN, V_V1UP, V_V1DN, V_TUP, V_TDN, V_RUP, V_RDN, MAXV };

// Statistics accumulated per strike (sum and sum of squares each).
enum Stat {
  ST_PRICE = 0, ST_DELTA, ST_GAMMA, ST_VEGA0, ST_VEGA1, ST_THETA, ST_RHO, ST_VANNA0, ST_VANNA1,
  ST_CHARM, NSTAT
};

// Fields staged per variant and sample for the strike loop.
enum Field { F_SPOT = 0, F_AUX, F_FLAGS, F_HT0, F_HT1, F_HT2, NFIELD };

constexpr int BLOCK = 256;      // lanes (= streams) per work unit
constexpr int STRIKE_CHUNK = 64;

template <typename R>
struct VarParams {
  R C[2][4];        // log-S coefficients per factor (init_cir_impl)
  R Kr[2][3];       // scheme constants in path precision
  double Kd[2][3];  // the same in FP64 (Poisson intensity, chi2 dof, scale)
  int nmax[2];      // NCI rows available to this variant's workspace
  R v0[2];
  R gdrift, gvol;      // zero-factor per-step drift and vol
  R gdrift_n, gvol_n;  // zero-factor European one-step: drift*n, vol*sqrt(n)
  R dt, total_time;
  double jmean, jexpm;  // Poisson mean of the jump span and exp(-mean)
  R jdrift;             // drift compensation over the span
  // FP64 copies of the log-S coefficients and drifts, for the maturity-bumped
  // variants that FP32 kernels step in FP64
  double Cd[2][4];
  double v0d[2];
  double jdrift_d, gdrift_nd, gvol_nd;
};

template <typename R>
struct Job {
  // model
  int nf, scheme[2], jump_kind, pathdep;
  R jmu, jdelta;            // Merton mu, delta
  R kp, keta1, keta2;       // Kou p, eta1, eta2
  // instrument / tracker
  int tracker, payoff, call, fixed_strike, barrier_up, barrier_in, rebate_timing, mode;
  double S0, logS0, mult_up, mult_dn, inv2h, invh2;
  R b_base, b_up, b_dn;               // log barriers (base, S0+h, S0-h)
  double init_sumlog, init_sumS;      // in-flight integrals past_avg * t_elapsed
  double hist_min_logS, hist_max_logS;
  double t_elapsed, maturity, rebate, rebate_rate, eps_scale;  // eps_scale = 1/(sqrt2 eps)
  // Greeks
  double inv2dv, inv2dT, inv2dr, cv[2], dfTu, dfTd, dfru, dfrd, erb_up, erb_dn;
  int nvega, derive_r;
  // variant groups: group g simulates var_list[first[g] .. first[g]+size[g])
  // in lockstep from one copy of the sample's stream state; group 0 holds the
  // base path and decides where the next sample starts.
  int n_groups, group_first[MAXV], group_size[MAXV], var_list[MAXV];
  int staged[MAXV];
  int n_fields;
  // sizes and buffers
  uint64_t M, S, steps;
  const uint64_t *streams;         // S x 4 stream states
  const void *nci[2];              // interleaved {y, dx*m} rows, R precision
  const double *nci_lgam[2];       // lgamma(d/2 + N) for N < nci_lg_n
  const double *nci_mean[2];       // per table row, the mean of its quantile map
  int nci_rows[2];
  int nci_lg_n;
  const void *nci_tail[2];         // upper-tail rows {q, hw*dq/dw} on w = -log(1-u); null = exact
  // FP64 copies of the main and tail rows for FP32 Full-mode jobs, read by
  // the maturity-bumped variants those kernels step in FP64 (null otherwise)
  const double *nci64[2];
  const double *nci_tail64[2];
  double tail_w0, tail_inv_hw;
  int tail_pts;
  int nci_grid;
  double nci_dx, nci_inv_dx;
  // u below this (the first grid interval, or 0 = never): rows N >= 2 are
  // inverted exactly instead of interpolated from the table's u = 0 node
  double nci_u_lo;
  // Rows above a variant's nmax (and u >= 0.999) that the reference inverts
  // exactly: rows N <= nci_rows_acc[f] at u in [nci_u_acc, nci_u_acc_hi) or
  // u >= 0.999 are read from the extended rows and the tail rows instead,
  // where the table agrees with the exact quantile to ~1e-11 relative (-1:
  // never; the extended-rows default already serves every row it holds)
  int nci_rows_acc[2];
  double nci_u_acc, nci_u_acc_hi;
  // randomized QMC (MYQL_STREAMS_SOBOL): Sobol' direction integers
  // V[k][d] (k = 0..31, QMC_DMAX coordinates each), the lower-6-bit table
  // lo[d][l], this replicate's digital shift (qmc_dims words), and the
  // coordinates a path may draw (0: pseudo-random streams)
  const uint32_t *qmc_V;
  const uint32_t *qmc_lo;
  const uint32_t *qmc_shift;
  int qmc_dims;
  const double *strikes;
  int nk, nkp;                     // strikes in this job, padded to a power of 2 (<= 64)
  double *partial;                 // per unit: 2*NSTAT x nk sums
  int first_slot;
  // Fused instruments: a job whose paths are identical to this job's (same
  // model, config, seed, maturity, Greek mode) is priced from this job's
  // staged samples instead of simulating them again. fuse_asian = 1: the
  // simulation also keeps the arithmetic-Asian state (this job a barrier,
  // the view an arithmetic Asian); view = 1 + that job's index (0: none);
  // view_accg: the view's per-lane accumulators, indexed by this launch's
  // unit like the job's own.
  int fuse_asian, view;
  double *view_accg;
  VarParams<R> var[MAXV];
};

struct Unit {
  int job;
  int blk;  // streams [blk*BLOCK, blk*BLOCK + BLOCK) of the job
};

}  // namespace myql
