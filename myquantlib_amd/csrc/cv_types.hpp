// Parameter blocks of the learned control-variate kernels (cv_kernel.hpp),
// shared with the host orchestration (engine.hip).
#pragma once
#include <stdint.h>

namespace myql {

constexpr int CV_H = 64;    // hidden width
constexpr int CV_IN = 5;    // features: t, log(S/S0), v1, v2, T (scaled, CvNorm)
constexpr int CV_OUT = 64;  // head: one output per strike of the strip
constexpr int CV_W1 = 0, CV_B1 = CV_W1 + CV_H * CV_IN, CV_W2 = CV_B1 + CV_H, CV_B2 = CV_W2 + CV_H * CV_H,
              CV_W3 = CV_B2 + CV_H, CV_B3 = CV_W3 + CV_H * CV_H, CV_W4 = CV_B3 + CV_H, CV_B4 = CV_W4 + CV_OUT * CV_H,
              CV_NPARAM = CV_B4 + CV_OUT;  // 12,864
static_assert(CV_NPARAM == 12864, "parameter layout");
constexpr int CV_MAXJ = 16;      // control points per path
constexpr int CV_REC = 6;        // recorded fields per (control point, path): x, v1, v2, A, V1, V2
constexpr int CV_PITCH = 72;     // FP16 row pitch of the LDS weight copies (bank spread)
constexpr int CV_TROWS = 32;     // rows per training block
constexpr int CV_NREG = 4;       // regressors per strike: M_A, M_V, M_J, D
constexpr int CV_RB_PTS = 2048;  // grid of the terminal-jump expectation tables
constexpr int CV_RB_TERMS = 40;  // jump counts n = 1..40 of those tables
constexpr int CV_MAXJOBS = 16;
// sums per strike written by the evaluation kernel: pricing (Y, Y^2, P, P^2,
// delta, delta^2, gamma, gamma^2), fitting (the 21 products of pairs of
// 1, P, M_A, M_V, M_J, D)
constexpr int CV_NS_PRICE = 8, CV_NS_FIT = 21;

// Feature scaling: f0 = t*ts, f1 = (x - log S0)*xs, f2 = v1*vs0, f3 = v2*vs1,
// f4 = T*ts.
struct CvNorm {
  float ts, xs, vs0, vs1, logS0, S0;
};

// One option set (a maturity's strikes) of a control-variate run.
struct CvSet {
  // recorded paths: rec[(j*CV_REC + c)*Ms + t], c = x, v1, v2 at control
  // point j, then A, V1, V2: the sums over block j of the stepper's exact
  // zero-mean increments (stock's own noise; each factor's variance noise)
  float *rec;
  double *ST;  // terminal spot (the plain pricer's value)
  double *Sm;  // terminal spot before the European's jump (null: no jumps)
  // paths, streams, and storage entries: the i-th path of stream s is
  // stored at t = i S + s, t < Ms = ceil(M / S) S (entries of streams with
  // one path fewer are unused)
  uint64_t M, S, Ms;
  int nj;
  int step_of[CV_MAXJ + 1];  // first step of each block, step_of[nj] = steps
  float t_of[CV_MAXJ];
  double disc_of[CV_MAXJ];   // e^{r (T - t_j)}
  int nf;
  float c2x[2];              // C2 of the log-S step per factor: the stock's loading on v'
  // NCI row means of the variance scheme's quantile map (record kernel)
  const double *nci_mean[2];
  int nci_rows[2];
  // instrument
  const double *strikes;     // device
  int nk, payoff, call, smooth, mode;
  double eps_scale, mult_up, mult_dn, inv2h, invh2;
  double T, df;
  // training-target normalisation per strike: (e^{-r(T-t)} P~ - mu_k) * inv_sd_k
  const float *mu, *inv_sd;
  // terminal-jump expectation E[P_k | S_T-] (Merton or Kou jumps; vanilla,
  // sharp and smoothed digital payoffs): p0 * P_k(S_T- e^{m0}) (no jump)
  // plus, on y = log S_T-,
  // the n >= 1 part rb[(i*nk + k)*2 + {0,1}] = {g, h g'} at grid point i
  // (cubic Hermite; strikes adjacent, so a path's lookups for all strikes
  // are contiguous); outside the grid the n >= 1 sum directly
  int rb_on;
  const double *rb;
  double rb_y0, rb_h, rb_inv_h;
  double rb_p0, rb_m0;
  // Merton: n-th term rb_p = Poisson weight, J_n ~ N(rb_m, rb_s^2).
  // Kou (rb_kou): the series coefficients kou_g[2 o + tilted][r] of
  // P(+-J > t) = sum_r G_r Q_r(e t) (orientation o = 0: up side first,
  // 1: mirrored; r < rb_n resp. rb_nd). Both laws: the n >= 1 part's mass
  // and E[e^J 1{n >= 1}] (J without m0) in kou_mass, kou_m1
  int rb_n, rb_kou, rb_nd;
  double kou_e1, kou_e2, kou_mass, kou_m1;
  double rb_p[CV_RB_TERMS], rb_m[CV_RB_TERMS], rb_s[CV_RB_TERMS];
  double kou_g[4][CV_RB_TERMS];
  // smoothed digitals: 16-point Gauss-Legendre nodes and weights on [-1, 1]
  // for the expectation over the payoff's transition window
  double gl_x[16], gl_w[16];
};

// Batched launches: task t owns blocks [first_block, next task's
// first_block) of the grid; jobs and sets are indices into the launch's
// device arrays (kernel arguments, so their fields load through the scalar
// cache).
struct CvPathTask {
  int job, set;
  unsigned first_block;
};
struct CvEvalTask {
  int set;
  unsigned first_block, n_waves;
  const double *coef;  // [CV_NREG][64] (pricing) or null (fitting)
  double *part;        // the waves' sums, [wave][64][ns]
  double *sums;        // [nk][ns]
};

}  // namespace myql
