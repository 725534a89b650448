// Learned martingale control variate (no counterpart in the reference, whose
// README.md:332 lists control variates as future work).
//
// A network u(t, X) -> R^64 gives, for every strike K_k of the strip at
// once, u_k(t, X) ~ E[e^{-r(T-t)} P_k | X_t] on the features
//   f = (t, log(S/S0), v1, v2, T) (each scaled, see CvNorm):
// an MLP 5 -> 64 -> 64 -> 64 -> 64 with SiLU activations whose last layer is
// a head over the strip's strikes. It is trained on the engine's own paths.
// Per path and strike the estimator is
//   Y_k = P_k - c_A M_A - c_V M_V - c_J M_J - c_D D
// with four zero-mean terms:
//   M_A = sum_j e^{r(T-t_j)} du_k/dx (t_j, X_j) A_j
//   M_V = sum_j e^{r(T-t_j)} sum_f (du_k/dx C2_f + du_k/dv_f)(t_j, X_j) V_f,j
//   M_J = e^{r(T-t_last)} du_k/dx (t_last, X_last) (e^J - 1)
//   D   = P_k - E[P_k | S_T-]   (Merton or Kou; vanilla and sharp digital
//                                payoffs)
// A_j and V_f,j are the block sums of the stepper's own increments whose
// conditional mean is exactly zero: the stock's independent noise
// sqrt(C3 (v + v')) G, and each factor's v' - E[v' | v] (closed form for NV
// and Exact; for NCI, v' - C2 m_N + 2 C2 (N - lambda) with m_N the mean of
// the table row's quantile map). The integrands are frozen at the control
// point t_j, so every term has zero mean whatever the network learned; the
// coefficients c are fitted per (maturity, strike) on the training paths,
// independent of the pricing paths, so the estimator stays unbiased.
//
// Kernels:
//   cv_paths_kernel     the engine's base path (same streams, partition and
//                       draws as mc_paths_kernel) recording state and the
//                       block increments at the control points, the
//                       terminal spot and the spot before the terminal jump
//   nci_rowmean_kernel  means of the NCI table rows' quantile maps
//   cv_rb_table_kernel  E[P_k | S_T-] of the n >= 1 jump counts on a log-spot
//                       grid
//   cv_moments_kernel   per-strike target scales of the training paths
//   cv_eval_kernel      per 16 paths and control point: forward-mode
//                       tangents of the hidden layers on FP16 MFMA (FP32
//                       accumulate) summed against the block increments,
//                       the heads once per path, then per (path, strike)
//                       the estimator (pricing) or the regressors' cross
//                       products (fitting), summed in FP64 by the strike's
//                       lane
//   cv_reduce_kernel    fixed-order sum of the waves' sums
//   cv_debug_rb_kernel  E[P_k | S_T-] at given spots (debug entry)
//   cv_train_kernel     FP32 forward/backward of 32 sampled rows per block
//   cv_adam_kernel      gradient reduction (fixed order) and Adam update
#pragma once
#include "cv_types.hpp"
#include "mlmc_kernel.hpp"

namespace myql {

// FP16 bit pattern of x (round to nearest even)
__device__ __forceinline__ short h16_bits(float x) { return __builtin_bit_cast(short, static_cast<_Float16>(x)); }

typedef short cv_v4s __attribute__((ext_vector_type(4)));
typedef float cv_v4f __attribute__((ext_vector_type(4)));

// logistic function with the hardware reciprocal (1 ulp; the network runs
// in FP16 on the matrix cores anyway)
__device__ __forceinline__ float sigmoidf(float a) { return __builtin_amdgcn_rcpf(1.0f + __expf(-a)); }

__device__ __forceinline__ double norm_cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }
__device__ __forceinline__ double norm_pdf(double x) { return 0.39894228040143267794 * exp(-0.5 * x * x); }

// Payoffs.hpp:14-69 (smoothed digitals in Greek modes, as the pricer)
__device__ __forceinline__ double cv_payoff(const CvSet &s, double S, double K) {
  const double diff = s.call ? S - K : K - S;
  if (s.payoff == 0) return diff > 0.0 ? diff : 0.0;
  const double w = s.smooth ? 0.5 * erfc(-diff * s.eps_scale) : (diff > 0.0 ? 1.0 : 0.0);
  return s.payoff == 1 ? w : S * w;
}

// E[P(S e^{J_n})] for J_n ~ N(m, s^2), s > 0, and its derivative in log S
// (vanilla, sharp cash and asset digitals)
__device__ __forceinline__ void rb_term(const CvSet &c, double S, double K, double m, double s, double &g,
                                        double &dg) {
  const double F = S * exp(m + 0.5 * s * s);
  const double d2 = (log(S / K) + m) / s, d1 = d2 + s;
  const double sg = c.call ? 1.0 : -1.0;
  if (c.payoff == 0) {
    g = sg * (F * norm_cdf(sg * d1) - K * norm_cdf(sg * d2));
    dg = sg * F * norm_cdf(sg * d1);
  } else if (c.payoff == 1) {
    g = norm_cdf(sg * d2);
    dg = sg * norm_pdf(d2) / s;
  } else {
    g = F * norm_cdf(sg * d1);
    dg = g + sg * F * norm_pdf(d1) / s;
  }
}

// Kou terminal jump J = A - B: A ~ Gamma(j, e1) the sum of j up jumps, B ~
// Gamma(m, e2) of m down jumps, j and m independent Poisson counts (a
// compound Poisson process thins into independent up and down parts).
// For integer shapes, conditioning on B and expanding (B + t)^i gives, at
// t >= 0, with a = e1 / (e1 + e2), b = e2 / (e1 + e2),
// Q_r(u) = e^{-u} sum_{k<r} u^k / k! (the upper regularised gamma function
// at integer shape) and u = e1 t:
//   P(A - B > t) = b^m sum_{l<j} C(m-1+l, l) a^l Q_{j-l}(u)
//   f_{A-B}(t)   = e1 b^m sum_{l<j} C(m-1+l, l) a^l e^{-u} u^{j-1-l} / (j-1-l)!
// (negative binomial weights: l down-events among the first events of the
// race of the two exponential clocks). Summed over j, m with weights
// wu_j cu^j, wd_m cd^m, the sums regroup by r = j - l into
//   P(t) = sum_{r>=1} G_r Q_r(u),  f(t) = e1 sum_{r>=1} G_r e^{-u} u^{r-1} / (r-1)!
// with G_r = sum_l NB_l wu_{l+r} cu^{l+r} and NB_l = sum_m wd_m cd^m
// C(m-1+l, l) a^l b^m independent of t (built on the host, kou_g); the
// (0, 0) term is zero at t >= 0.
__device__ __forceinline__ void kou_upper(const double *__restrict__ G, int n, double e1, double t, double &P,
                                          double &f) {
  const double u = e1 * t;
  double term = exp(-u), q = 0.0;
  P = 0.0;
  f = 0.0;
  for (int r = 1; r < n; ++r) {
    q += term;  // Q_r(u)
    f += G[r] * term;
    P += G[r] * q;
    term *= u / static_cast<double>(r);
  }
  f *= e1;
}

// n >= 1 part of P(J > t) and of the density of J at t (tilted = false), or
// of E[e^J 1{J > t}] (tilted = true: e^{A - B} turns Gamma(j, e1) into
// (e1 / (e1 - 1))^j Gamma(j, e1 - 1) and Gamma(m, e2) into
// (e2 / (e2 + 1))^m Gamma(m, e2 + 1)); t < 0 from the mirrored sum,
// P(A - B > t) = 1 - P(B - A >= -t)
__device__ __forceinline__ void kou_tail(const CvSet &c, bool tilted, double t, double &P, double &f) {
  const int o = tilted ? 1 : 0;
  if (t >= 0.0) {
    kou_upper(c.kou_g[o], c.rb_n, tilted ? c.kou_e1 - 1.0 : c.kou_e1, t, P, f);
  } else {
    double Pm;
    kou_upper(c.kou_g[2 + o], c.rb_nd, tilted ? c.kou_e2 + 1.0 : c.kou_e2, -t, Pm, f);
    P = (tilted ? c.kou_m1 : c.kou_mass) - Pm;
  }
}

// n >= 1 part of E[P(S e^{m0 + J})] under Kou jumps and its derivative in
// log S, for x = log(S / K) + m0: with T0 = P(J > -x), T1 = E[e^J 1{J > -x}]
// and f the density of J at -x, the call is K (e^x T1 - T0) with derivative
// K e^x T1, the cash digital T0 with derivative f, the asset digital
// K e^x T1 with derivative K e^x T1 + K f; puts by parity with the n >= 1
// mass and E[e^J 1{n >= 1}]
__device__ __noinline__ void kou_sum(const CvSet &c, double S, double K, double &g, double &dg) {
  const double x = log(S / K) + c.rb_m0;
  double T0, f, T1, f1;
  kou_tail(c, false, -x, T0, f);
  kou_tail(c, true, -x, T1, f1);
  const double F = S * exp(c.rb_m0);
  if (c.payoff == 0) {
    if (c.call) {
      g = F * T1 - K * T0;
      dg = F * T1;
    } else {
      g = K * (c.kou_mass - T0) - F * (c.kou_m1 - T1);
      dg = -F * (c.kou_m1 - T1);
    }
  } else if (c.payoff == 1) {
    g = c.call ? T0 : c.kou_mass - T0;
    dg = c.call ? f : -f;
  } else {
    g = c.call ? F * T1 : F * (c.kou_m1 - T1);
    dg = c.call ? F * T1 + K * f : F * (c.kou_m1 - T1) - K * f;
  }
}

// n >= 1 part of P(J > t) (tilted = false) or E[e^J 1{J > t}] (tilted =
// true), J the jumps' sum without the compensation, and the density of J
// at t (n >= 1 part): Kou in closed form (kou_tail); Merton from the normal
// n-jump sums N(n mu, n delta^2)
__device__ void jump_tail(const CvSet &c, bool tilted, double t, double &P, double &f) {
  if (c.rb_kou) {
    kou_tail(c, tilted, t, P, f);
    return;
  }
  P = 0.0;
  f = 0.0;
  for (int n = 0; n < c.rb_n; ++n) {
    const double m = c.rb_m[n] - c.rb_m0, s = c.rb_s[n];
    if (tilted) P += c.rb_p[n] * exp(m + 0.5 * s * s) * norm_cdf((m + s * s - t) / s);
    else P += c.rb_p[n] * norm_cdf((m - t) / s);
    f += c.rb_p[n] * norm_pdf((t - m) / s) / s;
  }
}

// n >= 1 part of E[P(S e^{m0 + J})] for an auto-smoothed digital,
// P(s) = Phi(+-(s - K) / eps) (times s for the asset digital), and its
// derivative in log S. Outside s in [K - 10 eps, K + 10 eps] the payoff is
// the sharp one to 1e-23, so those parts come from the tails in closed form;
// inside, Gauss-Legendre on 16 panels per side of the density's break at
// J = 0 integrates the smoothed payoff against the density of J.
__device__ __noinline__ void smooth_sum(const CvSet &c, double S, double K, double &g, double &dg) {
  const double F = S * exp(c.rb_m0), eps = 0.70710678118654752440 / c.eps_scale;
  const double sg = c.call ? 1.0 : -1.0;
  const double sl = K - 10.0 * eps, sh = K + 10.0 * eps;
  const double zh = log(sh / F), zl = sl > 0.0 ? log(sl / F) : zh - 60.0;
  double P, f;
  const bool asset = c.payoff == 2;
  const double mass = asset ? c.kou_m1 : c.kou_mass;
  if (c.call) {
    jump_tail(c, asset, zh, P, f);
    g = asset ? F * P : P;
  } else {
    jump_tail(c, asset, zl, P, f);
    g = asset ? F * (mass - P) : mass - P;
  }
  dg = asset ? g : 0.0;
  // the window, split at J = 0 where the Kou density breaks
  double e[3] = {zl, zh, zh};
  int ne = 1;
  if (zl < 0.0 && 0.0 < zh) {
    e[1] = 0.0;
    ne = 2;
  }
  constexpr int NP = 16;
  for (int q = 0; q < ne; ++q) {
    const double a = e[q], h = (e[q + 1] - e[q]) / NP;
    for (int pnl = 0; pnl < NP; ++pnl) {
      const double mid = a + (pnl + 0.5) * h;
      for (int i = 0; i < 16; ++i) {
        const double z = mid + 0.5 * h * c.gl_x[i], wq = 0.5 * h * c.gl_w[i];
        double Pz, fz;
        jump_tail(c, false, z, Pz, fz);
        const double s = F * exp(z), d = sg * (s - K) / eps;
        const double w = norm_cdf(d), dw = sg * norm_pdf(d) * s / eps;  // dw / dlog s
        g += wq * fz * (asset ? s * w : w);
        dg += wq * fz * (asset ? s * w + s * dw : dw);
      }
    }
  }
}

// n >= 1 part of E[P_k | S_T- = S] and its derivative in log S, summed
// directly
__device__ __forceinline__ void rb_sum(const CvSet &c, double S, double K, double &g, double &dg) {
  if (c.smooth) {
    smooth_sum(c, S, K, g, dg);
    return;
  }
  if (c.rb_kou) {
    kou_sum(c, S, K, g, dg);
    return;
  }
  g = 0.0;
  dg = 0.0;
  for (int n = 0; n < c.rb_n; ++n) {
    double a, b;
    rb_term(c, S, K, c.rb_m[n], c.rb_s[n], a, b);
    g += c.rb_p[n] * a;
    dg += c.rb_p[n] * b;
  }
}

// E[P_k | S_T- = Sm]: the no-jump term exactly, the rest from the table (or
// directly off the grid); lSm = log Sm, Sm0 = Sm e^{m0}
__device__ __forceinline__ double rb_expect(const CvSet &c, int k, double Sm, double lSm, double Sm0, double K) {
  const double e0 = c.rb_p0 * cv_payoff(c, Sm0, K);
  const double tp = (lSm - c.rb_y0) * c.rb_inv_h;
  double g;
  if (tp >= 0.0 && tp < static_cast<double>(CV_RB_PTS - 1)) {
    const int i = static_cast<int>(tp);
    const double t = tp - static_cast<double>(i), t2 = t * t, t3 = t2 * t;
    const double *a = c.rb + 2 * (static_cast<size_t>(i) * c.nk + k), *b = a + 2 * c.nk;
    g = (2.0 * t3 - 3.0 * t2 + 1.0) * a[0] + (t3 - 2.0 * t2 + t) * a[1] + (-2.0 * t3 + 3.0 * t2) * b[0] +
        (t3 - t2) * b[1];
  } else {
    double dg;
    rb_sum(c, Sm, K, g, dg);
  }
  return e0 + g;
}

// storage index t holds a path (the i-th of stream s, t = i S + s)
__device__ __forceinline__ bool cv_valid(const CvSet &c, uint64_t t) {
  const uint64_t i = t / c.S, s = t - i * c.S;
  return i < c.M / c.S + (s < c.M % c.S ? 1 : 0);
}

// ---------------------------------------------------------------------------
// recording path kernel
// ---------------------------------------------------------------------------
// task of this block in a batched launch
template <typename T>
__device__ __forceinline__ int cv_task_of(const T *__restrict__ tasks, int n) {
  int t = 0;
  while (t + 1 < n && blockIdx.x >= tasks[t + 1].first_block) ++t;
  return __builtin_amdgcn_readfirstlane(t);
}

// One thread per stream of each task's job, the tasks' blocks in one grid.
template <typename R, int NF>
__global__ void __launch_bounds__(BLOCK) cv_paths_kernel(const Job<R> *__restrict__ jobs,
                                                         const CvSet *__restrict__ sets,
                                                         const CvPathTask *__restrict__ tasks, int n_tasks) {
  constexpr int NFX = NF > 0 ? NF : 1;
  const int ti = cv_task_of(tasks, n_tasks);
  const Job<R> *__restrict__ J = jobs + __builtin_amdgcn_readfirstlane(tasks[ti].job);
  const CvSet &set = sets[__builtin_amdgcn_readfirstlane(tasks[ti].set)];
  const uint64_t S = J->S, M = J->M;
  const uint64_t s = static_cast<uint64_t>(blockIdx.x - tasks[ti].first_block) * BLOCK + threadIdx.x;
  if (s >= S) return;
  float *__restrict__ rec_out = set.rec;
  const int nj = set.nj;
  const int *__restrict__ nci_rows = set.nci_rows;
  const double *__restrict__ nci_mean[2] = {set.nci_mean[0], set.nci_mean[1]};
  Rng g;
  g.s0 = J->streams[4 * s + 0];
  g.s1 = J->streams[4 * s + 1];
  g.s2 = J->streams[4 * s + 2];
  g.s3 = J->streams[4 * s + 3];
  const uint64_t q = M / S, rem = M % S;
  const uint64_t count = q + (s < rem ? 1 : 0);
  const uint64_t Ms = set.Ms;
  const VarParams<R> *__restrict__ P = &J->var[V_BASE];
  const VarParams<R> *__restrict__ PP[1] = {P};
  const unsigned steps = static_cast<unsigned>(J->steps);
  // the i-th path of stream s (its i-th sample, the same draws as the
  // plain pricer's) is stored at t = i S + s, so the lanes of a wave write
  // consecutive entries
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t p = i * S + s;
    // the log-return or the log-price, as the pricing kernels carry it
    // (recorded as the log-price)
    constexpr bool LR = log_return_paths<R, false>();
    const double xrec0 = LR ? J->logS0 : 0.0;
    R x = LR ? R(0) : static_cast<R>(J->logS0);
    R v[NFX];
#pragma unroll
    for (int f = 0; f < NFX; ++f) v[f] = P->v0[f];
    int j = 0;
    double A = 0.0, V[2] = {0.0, 0.0};
    auto put = [&](int jj, int c, double val) {
      rec_out[(static_cast<size_t>(jj) * CV_REC + c) * Ms + p] = static_cast<float>(val);
    };
    if constexpr (NF == 0) {
      // zero-factor European: one step to maturity (ASVJstepper.hpp:145-161),
      // one control point at t = 0
      put(0, 0, xrec0 + static_cast<double>(x));
      put(0, 1, 0.0);
      put(0, 2, 0.0);
      const R G = normcdfinv<R>(uniform(g));
      x += P->gdrift_n + P->gvol_n * G;
      A = static_cast<double>(P->gvol_n * G);
    } else {
      for (unsigned step = 0; step < steps; ++step) {
        if (j < nj && step == static_cast<unsigned>(set.step_of[j])) {
          if (j > 0) {  // close the previous block
            put(j - 1, 3, A);
            put(j - 1, 4, V[0]);
            put(j - 1, 5, V[1]);
          }
          put(j, 0, xrec0 + static_cast<double>(x));
          put(j, 1, v[0]);
          put(j, 2, NF > 1 ? static_cast<double>(v[NFX - 1]) : 0.0);
          A = 0.0;
          V[0] = V[1] = 0.0;
          ++j;
        }
        // one step of ASVJStepper::multi_step with the draws the engine makes
        const R G = normcdfinv<R>(uniform(g));
        R vn[NFX];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const double vd = static_cast<double>(v[f]);
          if (J->scheme[f] == 1) {
            const R Z = normcdfinv<R>(uniform(g));
            vn[f] = nv_step<R, NF>(P, f, v[f], Z);
            // E[v'|v] = K0 (K0 v + K1 + K2^2) + K1 of the NV step
            const double K0 = P->Kr[f][0], K1 = P->Kr[f][1], K2 = P->Kr[f][2];
            V[f] += static_cast<double>(vn[f]) - (K0 * (K0 * vd + K1 + K2 * K2) + K1);
          } else if (J->scheme[f] == 2) {
            int N = 0;
            vn[f] = evolve_one<R, NFX>(J, P, f, v[f], g, &N);
            const double lam = P->Kd[f][0] * vd;
            const double mN = N < nci_rows[f] ? nci_mean[f][N] : P->Kd[f][1] + 2.0 * N;
            V[f] += static_cast<double>(vn[f]) - P->Kd[f][2] * mN + 2.0 * P->Kd[f][2] * (N - lam);
          } else {
            vn[f] = evolve_one<R, NFX>(J, P, f, v[f], g);
            // E[v'|v] = C2 (alpha + C0 v), the exact CIR mean
            V[f] += static_cast<double>(vn[f]) - P->Kd[f][2] * (P->Kd[f][1] + P->Kd[f][0] * vd);
          }
        }
        R a = 0, d = 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          a += P->C[f][3] * (v[f] + vn[f]);
          d += P->C[f][0] + P->C[f][1] * v[f] + P->C[f][2] * vn[f];
        }
        const R sG = dsqrt<R>(a) * G;
        x += d + sG;
        A += static_cast<double>(sG);
#pragma unroll
        for (int f = 0; f < NF; ++f) v[f] = vn[f];
      }
    }
    put(nj - 1, 3, A);
    put(nj - 1, 4, V[0]);
    put(nj - 1, 5, V[1]);
    const R xm = x;
    R xs[1] = {x};
    euro_jump<R, 1>(J, PP, xs, g);
    const R S0r = LR ? static_cast<R>(J->S0) : R(1);
    set.ST[p] = static_cast<double>(S0r * dexp<R>(xs[0]));
    if (set.Sm) set.Sm[p] = static_cast<double>(S0r * dexp<R>(xm));
  }
}

// ---------------------------------------------------------------------------
// NCI table row means: m_N = integral over u in [0, 1) of the quantile map the
// variance step applies to row N (engine default semantics): the exact chi2
// quantile on [0, dx) for N >= 2, the cubic Hermite pieces on [dx, 0.999) (all
// of [0, 0.999) for N < 2), the tail above 0.999 (exact mean there; the tail
// rows agree with the exact quantile to ~1e-11). With k = d + 2N degrees of
// freedom, the exact pieces are E[X 1{X < q}] = k P(k/2 + 1, q/2) and its
// complement.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double reg_gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  if (x < a + 1.0) return igam_series(a, x, lgamma(a + 1.0));
  return 1.0 - igamc_cf(a, x, lgamma(a));
}

template <typename R>
__global__ void __launch_bounds__(256) nci_rowmean_kernel(const R *__restrict__ tab, int rows, int grid, double dx,
                                                          double d, double *__restrict__ mean) {
  const int N = blockIdx.x;
  if (N >= rows) return;
  __shared__ double red[256];
  const R *row = tab + 2 * static_cast<size_t>(N) * grid;
  const int i0 = N >= 2 ? 1 : 0;
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < grid - 1; i += blockDim.x) {
    const double y0 = row[2 * i], m0 = row[2 * i + 1], y1 = row[2 * i + 2], m1 = row[2 * i + 3];
    s += 0.5 * (y0 + y1) + (m0 - m1) * (1.0 / 12.0);
  }
  red[threadIdx.x] = s * dx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double k = d + 2.0 * N, a = 0.5 * k + 1.0;
    double m = red[0];
    if (N >= 2) m += k * reg_gamma_p(a, 0.5 * chi2_quantile(k, dx));
    m += k * (1.0 - reg_gamma_p(a, 0.5 * chi2_quantile(k, 0.999)));
    mean[N] = m;
  }
}

// E[P_k | S_T-] of the n >= 1 jump counts on the log-spot grid
__global__ void cv_rb_table_kernel(CvSet set, double *__restrict__ tab) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= set.nk * CV_RB_PTS) return;
  const int i = idx / set.nk, k = idx % set.nk;
  const double y = set.rb_y0 + i * set.rb_h;
  double g, dg;
  rb_sum(set, exp(y), set.strikes[k], g, dg);
  tab[2 * idx] = g;
  tab[2 * idx + 1] = set.rb_h * dg;
}

// E[P_k | S_T- = spots[i]] per (spot i, strike k): from the table as the
// estimator reads it (tab), and summed directly (dir; with its derivative in
// log S of the n >= 1 part, ddir)
__global__ void cv_debug_rb_kernel(CvSet set, const double *__restrict__ spots, int64_t n, double *__restrict__ tab,
                                   double *__restrict__ dir, double *__restrict__ ddir) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * set.nk) return;
  const int64_t i = idx / set.nk;
  const int k = static_cast<int>(idx % set.nk);
  const double S = spots[i], K = set.strikes[k], S0 = S * exp(set.rb_m0);
  tab[idx] = rb_expect(set, k, S, log(S), S0, K);
  double g, dg;
  rb_sum(set, S, K, g, dg);
  dir[idx] = set.rb_p0 * cv_payoff(set, S0, K) + g;
  ddir[idx] = dg;
}

// discounted (terminal-jump-expected) payoff of path p, strike k: the
// training target before normalisation
__device__ __forceinline__ double cv_target(const CvSet &c, uint64_t p, int k) {
  const double K = c.strikes[k];
  if (c.rb_on) {
    const double Sm = c.Sm[p];
    return rb_expect(c, k, Sm, log(Sm), Sm * exp(c.rb_m0), K);
  }
  return cv_payoff(c, c.ST[p], K);
}

// sums of df * target and its square per (strike, block); grid (blocks, nk)
__global__ void __launch_bounds__(256) cv_moments_kernel(CvSet set, double *__restrict__ part) {
  __shared__ double r0[256], r1[256];
  const int k = blockIdx.y;
  double s = 0.0, s2 = 0.0;
  for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < set.Ms;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (!cv_valid(set, p)) continue;
    const double y = set.df * cv_target(set, p, k);
    s += y;
    s2 += y * y;
  }
  r0[threadIdx.x] = s;
  r1[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] += r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (static_cast<size_t>(k) * gridDim.x + blockIdx.x)] = r0[0];
    part[2 * (static_cast<size_t>(k) * gridDim.x + blockIdx.x) + 1] = r1[0];
  }
}

// ---------------------------------------------------------------------------
// network weights in LDS for the MFMA evaluation
// ---------------------------------------------------------------------------
// All products run on v_mfma_f32_16x16x16_f16 in the transposed form
// out^T = W x^T (16 units x 16 rows per tile): the accumulator of output
// tile m (lane l holds units 16m + 4(l/16) + 0..3 of row l % 16) is exactly
// the B fragment of input chunk m of the next product, so activations never
// leave the lane. Layer 1 is a product of W1 extended by the bias column
// against (f, 1); the biases of layers 2 and 3 seed the accumulators, and
// the columns of W1 seed the tangents. hd[0..3): the head's rows W4[k]
// (fitting, checks) or, for pricing, W4[k] scaled by xs c_A,k, xs c_V,k and
// xs c_J,k, so that a head product is the coefficient-weighted tangent.
// The value pass runs in base-2 units: pre-activations y' = y log2(e) (W1,
// b1, b2, b3 scaled by log2(e)), so the sigmoid is 1 / (1 + 2^-y') with no
// scaling multiply, and the activations it passes on are h' = h log2(e),
// which W2 and W3 unscaled turn into the next y' (log2(e) ln(2) = 1).
constexpr int CV_P1 = 20;  // FP16 pitch of the 16-column matrix W1
constexpr float kCvLog2e = 1.44269504088896341f, kCvLn2 = 0.69314718055994531f;
struct CvLds {
  short w2[CV_H * CV_PITCH], w3[CV_H * CV_PITCH];  // [out][in]
  short hd[3][CV_OUT * CV_PITCH];
  short w1[CV_H * CV_P1];                 // [unit][W1 0..4, b1, 0...]
  alignas(16) float bias[2][CV_H];        // b2, b3
  alignas(16) float w1c[3][CV_H];         // columns 1..3 of W1
};

__device__ __forceinline__ void cv_load_lds(CvLds &L, const float *__restrict__ prm, const double *__restrict__ coef,
                                            float xs, int nk) {
  for (int i = threadIdx.x; i < CV_H * CV_H; i += blockDim.x) {
    const int n = i / CV_H, k = i % CV_H;
    L.w2[n * CV_PITCH + k] = h16_bits(prm[CV_W2 + i]);
    L.w3[n * CV_PITCH + k] = h16_bits(prm[CV_W3 + i]);
    if (!coef) {
      L.hd[0][n * CV_PITCH + k] = h16_bits(prm[CV_W4 + i]);
    } else {
      // row n of the head is strike n
#pragma unroll
      for (int h = 0; h < 3; ++h) {
        const float sc = n < nk ? static_cast<float>(xs * coef[h * CV_OUT + n]) : 0.0f;
        L.hd[h][n * CV_PITCH + k] = h16_bits(sc * prm[CV_W4 + i]);
      }
    }
  }
  for (int i = threadIdx.x; i < CV_H * 16; i += blockDim.x) {
    const int n = i / 16, k = i % 16;
    const float w = k < CV_IN ? prm[CV_W1 + n * CV_IN + k] : (k == CV_IN ? prm[CV_B1 + n] : 0.0f);
    L.w1[n * CV_P1 + k] = h16_bits(w * kCvLog2e);
  }
  for (int i = threadIdx.x; i < CV_H; i += blockDim.x) {
    L.bias[0][i] = prm[CV_B2 + i] * kCvLog2e;
    L.bias[1][i] = prm[CV_B3 + i] * kCvLog2e;
#pragma unroll
    for (int c = 0; c < 3; ++c) L.w1c[c][i] = prm[CV_W1 + i * CV_IN + 1 + c];
  }
}

typedef cv_v4f cv_h4;  // SiLU derivatives of one output tile

// Unit held by lane l in element e of output tile m.
__device__ __forceinline__ int cv_unit(int m, int e, int l) { return 16 * m + 4 * (l >> 4) + e; }

// A zero the compiler cannot see through: LDS weight reads indexed with it
// are re-issued on every product instead of being hoisted out of the
// caller's path loop into registers.
__device__ __forceinline__ int cv_opaque_zero() {
  int z = 0;
  asm volatile("" : "+s"(z));
  return z;
}

typedef _Float16 cv_v4h __attribute__((ext_vector_type(4)));

__device__ __forceinline__ cv_v4f cv_mfma(cv_v4s a, cv_v4s b, cv_v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(cv_v4h, a), __builtin_bit_cast(cv_v4h, b), c, 0, 0,
                                                0);
}

// out^T = W x^T for 64 outputs and 64 inputs: acc[m] += W[16m.., :] x^T
__device__ __forceinline__ void cv_layer(const short *__restrict__ W, const cv_v4s (&B)[4], cv_v4f (&acc)[4], int l) {
  const int z = cv_opaque_zero();
  const short *base = W + (l & 15) * CV_PITCH + 4 * (l >> 4) + z;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      acc[m] = cv_mfma(*reinterpret_cast<const cv_v4s *>(base + 16 * m * CV_PITCH + 16 * s), B[s], acc[m]);
}

// the lane's units 16m + 4(l/16) + 0..3 of a 64-float vector in LDS
__device__ __forceinline__ cv_v4f cv_units4(const float *__restrict__ v, int m, int l) {
  return *reinterpret_cast<const cv_v4f *>(v + 16 * m + 4 * (l >> 4) + cv_opaque_zero());
}

typedef unsigned cv_v2u __attribute__((ext_vector_type(2)));
typedef __fp16 cv_v2h __attribute__((ext_vector_type(2)));

// four FP32 values as a B fragment: two packed conversions (round toward
// zero), one instruction per pair
__device__ __forceinline__ cv_v4s cv_pack4(cv_v4f x) {
  const cv_v2h lo = __builtin_amdgcn_cvt_pkrtz(x[0], x[1]), hi = __builtin_amdgcn_cvt_pkrtz(x[2], x[3]);
  return __builtin_bit_cast(cv_v4s, cv_v2u{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)});
}

__device__ __forceinline__ void cv_to_h16(const cv_v4f (&x)[4], cv_v4s (&B)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) B[m] = cv_pack4(x[m]);
}

// Value pass of the row held by lanes (l % 16) + 16 q: the SiLU derivatives
// of the three hidden layers, dd[layer][m][e] for unit cv_unit(m, e, l).
// With sigma = 1 / (1 + 2^-y'), h' = y' sigma: dh/dy = sigma + h (1 - sigma)
// = sigma + ln(2) h' (1 - sigma).
__device__ __forceinline__ void cv_value(const CvLds &L, const float (&f)[CV_IN], cv_h4 (&dd)[3][4], int l) {
  // layer 1: (f0..f4, 1) against W1 extended by b1
  cv_v4s Bf = {0, 0, 0, 0};
  if ((l >> 4) == 0) {
    Bf[0] = h16_bits(f[0]);
    Bf[1] = h16_bits(f[1]);
    Bf[2] = h16_bits(f[2]);
    Bf[3] = h16_bits(f[3]);
  } else if ((l >> 4) == 1) {
    Bf[0] = h16_bits(f[4]);
    Bf[1] = h16_bits(1.0f);
  }
  cv_v4f acc[4];
  {
    const int z = cv_opaque_zero();
    const short *base = L.w1 + (l & 15) * CV_P1 + 4 * (l >> 4) + z;
#pragma unroll
    for (int m = 0; m < 4; ++m)
      acc[m] = cv_mfma(*reinterpret_cast<const cv_v4s *>(base + 16 * m * CV_P1), Bf, cv_v4f{0.f, 0.f, 0.f, 0.f});
  }
  cv_v4s Bh[4];
#pragma unroll
  for (int layer = 0; layer < 3; ++layer) {
    if (layer > 0) {
#pragma unroll
      for (int m = 0; m < 4; ++m) acc[m] = cv_units4(L.bias[layer - 1], m, l);
      cv_layer(layer == 1 ? L.w2 : L.w3, Bh, acc, l);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      cv_v4f h;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float y = acc[m][e];
        const float sg = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-y));
        h[e] = y * sg;
        dd[layer][m][e] = (h[e] - h[e] * sg) * kCvLn2 + sg;
      }
      if (layer < 2) Bh[m] = cv_pack4(h);
    }
  }
}

// Forward-mode tangent along feature i (1..3) through the hidden layers:
// the last hidden layer's tangent t (FP32, accumulator layout).
__device__ __forceinline__ void cv_tangent(const CvLds &L, const cv_h4 (&dd)[3][4], int i, cv_v4f (&t)[4], int l) {
  cv_v4s Bt[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) Bt[m] = cv_pack4(cv_units4(L.w1c[i - 1], m, l) * dd[0][m]);
#pragma unroll
  for (int layer = 1; layer < 3; ++layer) {
    cv_v4f acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = cv_v4f{0.f, 0.f, 0.f, 0.f};
    cv_layer(layer == 1 ? L.w2 : L.w3, Bt, acc, l);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const cv_v4f v = acc[m] * dd[layer][m];
      if (layer == 1) Bt[m] = cv_pack4(v);
      else t[m] = v;
    }
  }
}

// head product: out[m] (+)= H[16m.., :] x^T of the FP32 hidden-layer vector x
template <bool ADD>
__device__ __forceinline__ void cv_head(const short *__restrict__ H, const cv_v4f (&x)[4], cv_v4f (&out)[4], int l) {
  cv_v4s B[4];
  cv_to_h16(x, B);
  if (!ADD)
#pragma unroll
    for (int m = 0; m < 4; ++m) out[m] = cv_v4f{0.f, 0.f, 0.f, 0.f};
  cv_layer(H, B, out, l);
}

// Orders a wave's LDS accesses (DS instructions of one wave execute in
// issue order; this keeps the compiler from moving them across).
__device__ __forceinline__ void cv_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Terminal data of the 16 paths of a wave's batch, for the per-strike stage.
struct CvPathLds {
  double ST[16], Sm[16], Sm0[16], hw[4][16];  // hw: cubic Hermite weights
  int cell[16];                               // table cell, -1: off the grid
  int valid[16];
};

// ---------------------------------------------------------------------------
// evaluation. Storage index t = i S + s holds the i-th path of stream s
// (lanes of a wave read consecutive streams); each wave takes 16 storage
// indices at a time. Per control point: the value pass, then each tangent
// through the hidden layers, accumulated (weighted by the block increments)
// into per-path FP32 hidden-layer sums TA (stock noise), TV (variance
// noise) and TJ (terminal jump). After the last point the heads turn them
// into the strikes' martingales: MODE 0/1 Z = c_A M_A + c_V M_V + c_J M_J
// (heads pre-scaled by the coefficients), MODE 2 the three regressors. The
// outputs pass through LDS to the layout lane = strike, and each lane
// accumulates its strike's sums over the wave's paths in FP64, in path
// order: MODE 0 the estimator Y = P - Z - c_D D; 1 the same with the plain
// delta and gamma of the path (spot scaled by (S0 +- h)/S0,
// European.hpp:40-74); 2 the products of pairs of (1, P, M_A, M_V, M_J, D)
// for the fit. part[(wave * 64 + k) * NS + s] receives the wave's sums.
// ---------------------------------------------------------------------------
template <int NF, int MODE>
__global__ void __launch_bounds__(BLOCK) cv_eval_kernel(const float *__restrict__ prm, CvNorm nm,
                                                        const CvSet *__restrict__ sets,
                                                        const CvEvalTask *__restrict__ tasks, int n_tasks) {
  constexpr int NS = MODE == 2 ? CV_NS_FIT : (MODE == 1 ? CV_NS_PRICE : 4);
  constexpr int NZ = MODE == 2 ? 3 : 1;
  constexpr int WAVES = BLOCK / 64, ZP = CV_OUT + 4;
  __shared__ CvLds L;
  __shared__ double sK[CV_OUT];
  __shared__ float sD[CV_OUT];
  __shared__ alignas(16) float sZ[WAVES][NZ * 16 * ZP];
  __shared__ CvPathLds sP[WAVES];
  const int ti = cv_task_of(tasks, n_tasks);
  const CvSet &set = sets[__builtin_amdgcn_readfirstlane(tasks[ti].set)];
  const double *__restrict__ coef = tasks[ti].coef;
  double *__restrict__ part = tasks[ti].part;
  cv_load_lds(L, prm, MODE == 2 ? nullptr : coef, nm.xs, set.nk);
  for (int i = threadIdx.x; i < CV_OUT; i += blockDim.x) {
    const bool live = i < set.nk;
    sK[i] = live ? set.strikes[i] : 0.0;
    sD[i] = (live && coef && MODE != 2) ? static_cast<float>(coef[3 * CV_OUT + i]) : 0.0f;
  }
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x / 64;
  const int r = l & 15;
  const uint64_t n_waves = tasks[ti].n_waves;
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x - tasks[ti].first_block) * WAVES + wv;
  const uint64_t Ms = set.Ms;
  const float c2x0 = set.c2x[0], c2x1 = set.c2x[1];
  // weights of the tangent sums: MODE 2 unscaled regressors (head W4, so
  // du/dx = xs W4 t, du/dv_f = vs_f W4 t_f); MODE 0/1 heads carry xs
  const float sx = MODE == 2 ? nm.xs : 1.0f;
  const float rv0 = MODE == 2 ? nm.vs0 : nm.vs0 / nm.xs, rv1 = MODE == 2 ? nm.vs1 : nm.vs1 / nm.xs;
  const float fT = static_cast<float>(set.T) * nm.ts;
  const int k = l;  // strike of the per-strike stage
  const bool k_live = k < set.nk;
  const double K = sK[k], cD = sD[k];
  float *zb = sZ[wv];
  CvPathLds &pd = sP[wv];
  double acc[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) acc[i] = 0.0;

  for (uint64_t p0 = wave * 16; p0 < Ms; p0 += n_waves * 16) {
    const uint64_t p = p0 + r;
    const bool valid = p < Ms && cv_valid(set, p);
    const uint64_t pp = p < Ms ? p : Ms - 1;
    cv_v4f TA[4], TV[4], TJ[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      TA[m] = cv_v4f{0.f, 0.f, 0.f, 0.f};
      TV[m] = cv_v4f{0.f, 0.f, 0.f, 0.f};
      TJ[m] = cv_v4f{0.f, 0.f, 0.f, 0.f};
    }
    const double ST = set.ST[pp];
    const double Sm = set.Sm ? set.Sm[pp] : ST;
    const float ej = set.Sm ? static_cast<float>(ST / Sm - 1.0) : 0.0f;
    for (int j = 0; j < set.nj; ++j) {
      const float *rec = set.rec + static_cast<size_t>(j) * CV_REC * Ms + pp;
      const float x = rec[0], v1 = rec[Ms], v2 = rec[2 * Ms], A = rec[3 * Ms], V1 = rec[4 * Ms],
                  V2 = NF > 1 ? rec[5 * Ms] : 0.0f;
      const float f[CV_IN] = {set.t_of[j] * nm.ts, (x - nm.logS0) * nm.xs, v1 * nm.vs0, v2 * nm.vs1, fT};
      const float dj = static_cast<float>(set.disc_of[j]);
      const float Cv = c2x0 * V1 + (NF > 1 ? c2x1 * V2 : 0.0f);
      const bool last = j == set.nj - 1;
      cv_h4 dd[3][4];
      cv_value(L, f, dd, l);
      cv_v4f t[4];
      // log-spot tangent: stock noise, the variance noise's stock part and
      // (last point) the terminal jump
      cv_tangent(L, dd, 1, t, l);
      const float wa = dj * sx * A, wv1 = dj * sx * Cv, wj = dj * sx * ej;
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          TA[m][e] += wa * t[m][e];
          TV[m][e] += wv1 * t[m][e];
          if (last) TJ[m][e] = wj * t[m][e];
        }
      // variance tangents (their feature scale relative to the log-spot's
      // folded into the weight)
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) {
        cv_tangent(L, dd, 2 + fi, t, l);
        const float w = dj * (fi == 0 ? rv0 * V1 : rv1 * V2);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int e = 0; e < 4; ++e) TV[m][e] += w * t[m][e];
      }
    }
    // terminal data of the batch's paths
    if (l < 16) {
      pd.ST[r] = ST;
      pd.Sm[r] = Sm;
      pd.valid[r] = valid ? 1 : 0;
      if (set.rb_on) {
        pd.Sm0[r] = Sm * exp(set.rb_m0);
        const double tp = (log(Sm) - set.rb_y0) * set.rb_inv_h;
        const bool on = tp >= 0.0 && tp < static_cast<double>(CV_RB_PTS - 1);
        const int i = on ? static_cast<int>(tp) : -1;
        const double u = on ? tp - static_cast<double>(i) : 0.0, u2 = u * u, u3 = u2 * u;
        pd.cell[r] = i;
        pd.hw[0][r] = 2.0 * u3 - 3.0 * u2 + 1.0;
        pd.hw[1][r] = u3 - 2.0 * u2 + u;
        pd.hw[2][r] = -2.0 * u3 + 3.0 * u2;
        pd.hw[3][r] = u3 - u2;
      }
    }
    // heads, then through LDS to lane = strike
#pragma unroll
    for (int q = 0; q < NZ; ++q) {
      cv_v4f o[4];
      if constexpr (MODE == 2) {
        cv_head<false>(L.hd[0], q == 0 ? TA : (q == 1 ? TV : TJ), o, l);
      } else {
        cv_head<false>(L.hd[0], TA, o, l);
        cv_head<true>(L.hd[1], TV, o, l);
        if (set.Sm) cv_head<true>(L.hd[2], TJ, o, l);
      }
#pragma unroll
      for (int m = 0; m < 4; ++m)
        *reinterpret_cast<cv_v4f *>(zb + (q * 16 + r) * ZP + 16 * m + 4 * (l >> 4)) = o[m];
    }
    cv_wave_sync();
    if (k_live) {
      for (int i = 0; i < 16; ++i) {
        if (!pd.valid[i]) continue;
        const double STi = pd.ST[i];
        const double P = cv_payoff(set, STi, K);
        double D = 0.0;
        if (set.rb_on) {
          const int cell = pd.cell[i];
          double g;
          if (cell >= 0) {
            const double *a = set.rb + 2 * (static_cast<size_t>(cell) * set.nk + k), *b = a + 2 * set.nk;
            g = pd.hw[0][i] * a[0] + pd.hw[1][i] * a[1] + pd.hw[2][i] * b[0] + pd.hw[3][i] * b[1];
          } else {
            double dg;
            rb_sum(set, pd.Sm[i], K, g, dg);
          }
          D = P - (set.rb_p0 * cv_payoff(set, pd.Sm0[i], K) + g);
        }
        if constexpr (MODE == 2) {
          const double v[6] = {1.0, P, zb[i * ZP + k], zb[(16 + i) * ZP + k], zb[(32 + i) * ZP + k], D};
          int q = 0;
#pragma unroll
          for (int u = 0; u < 6; ++u)
#pragma unroll
            for (int w = u; w < 6; ++w) acc[q++] += v[u] * v[w];
        } else {
          const double Y = P - static_cast<double>(zb[i * ZP + k]) - cD * D;
          acc[0] += Y;
          acc[1] += Y * Y;
          acc[2] += P;
          acc[3] += P * P;
          if constexpr (MODE == 1) {
            const double pu = cv_payoff(set, STi * set.mult_up, K), pdn = cv_payoff(set, STi * set.mult_dn, K);
            const double d = (pu - pdn) * set.inv2h, gm = (pu - 2.0 * P + pdn) * set.invh2;
            acc[4] += d;
            acc[5] += d * d;
            acc[6] += gm;
            acc[7] += gm * gm;
          }
        }
      }
    }
    cv_wave_sync();
  }
  if (k_live)
    for (int i = 0; i < NS; ++i) part[(wave * CV_OUT + k) * NS + i] = acc[i];
}

// sums[k][s] = sum over waves of part[w][k][s] of task blockIdx.y: one
// block per (k, s), strided partial sums then a tree (fixed order)
__global__ void __launch_bounds__(256) cv_reduce_kernel(const CvSet *__restrict__ sets,
                                                        const CvEvalTask *__restrict__ tasks, int ns) {
  __shared__ double red[256];
  const CvEvalTask &T = tasks[blockIdx.y];
  if (static_cast<int>(blockIdx.x) >= sets[T.set].nk * ns) return;
  const double *__restrict__ part = T.part;
  const int n_waves = static_cast<int>(T.n_waves);
  const int k = blockIdx.x / ns, c = blockIdx.x % ns;
  double s = 0.0;
  for (int w = threadIdx.x; w < n_waves; w += 256) s += part[(static_cast<size_t>(w) * CV_OUT + k) * ns + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) T.sums[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t cv_hash(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct CvTrainLds {
  float f[CV_TROWS][CV_IN];
  float h1[CV_TROWS][CV_H + 1], h2[CV_TROWS][CV_H + 1], h3[CV_TROWS][CV_H + 1];
  float d1[CV_TROWS][CV_H + 1], d2[CV_TROWS][CV_H + 1], d3[CV_TROWS][CV_H + 1];
  float dy[CV_TROWS][CV_OUT + 1];
  float loss[CV_TROWS][8];
  int set[CV_TROWS];
  unsigned long long path[CV_TROWS];
  int cp[CV_TROWS];
};

// Rows: a random (set, path, control point) per row of each minibatch, from
// a counter hash of (seed, step, row). Targets: the normalised discounted
// (terminal-jump-expected) payoff of each strike, masked beyond the set's
// strike count. Writes the block's parameter gradient of the mean squared
// error over the minibatch's rows and live strikes, and its loss. Weights
// are read from global memory (L2 resident).
__global__ void __launch_bounds__(BLOCK) cv_train_kernel(const float *__restrict__ prm, const CvSet *__restrict__ sets,
                                                         int n_sets, CvNorm nm, uint64_t seed,
                                                         const int *__restrict__ step, int batch, int nk_all,
                                                         float *__restrict__ gpart, float *__restrict__ lpart,
                                                         const float *__restrict__ fixed_f,
                                                         const float *__restrict__ fixed_t) {
  __shared__ CvTrainLds L;
  const int tid = threadIdx.x;
  const int r = tid % CV_TROWS, q8 = tid / CV_TROWS;  // row, slice of 8 units
  const int row = blockIdx.x * CV_TROWS + r;
  if (q8 == 0) {
    float f[CV_IN];
    if (fixed_f) {
      for (int i = 0; i < CV_IN; ++i) f[i] = fixed_f[row * CV_IN + i];
      L.set[r] = -1;
    } else {
      const uint64_t h = cv_hash(seed ^ cv_hash(static_cast<uint64_t>(*step) * 0x100000001ull + row));
      const int si = static_cast<int>(h % static_cast<uint64_t>(n_sets));
      const CvSet &S = sets[si];
      const uint64_t h2 = cv_hash(h);
      uint64_t p = h2 % S.Ms;
      if (!cv_valid(S, p)) p -= S.S;  // the stream's previous path
      const int j = static_cast<int>((h2 >> 40) % static_cast<uint64_t>(S.nj));
      f[0] = S.t_of[j] * nm.ts;
      f[1] = (S.rec[(static_cast<size_t>(j) * CV_REC + 0) * S.Ms + p] - nm.logS0) * nm.xs;
      f[2] = S.rec[(static_cast<size_t>(j) * CV_REC + 1) * S.Ms + p] * nm.vs0;
      f[3] = S.rec[(static_cast<size_t>(j) * CV_REC + 2) * S.Ms + p] * nm.vs1;
      f[4] = static_cast<float>(S.T) * nm.ts;
      L.set[r] = si;
      L.path[r] = p;
      L.cp[r] = j;
    }
    for (int i = 0; i < CV_IN; ++i) L.f[r][i] = f[i];
  }
  __syncthreads();
  const float inv_n = 1.0f / (static_cast<float>(batch) * static_cast<float>(nk_all));
  // forward
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = prm[CV_B1 + n];
    for (int i = 0; i < CV_IN; ++i) z += prm[CV_W1 + n * CV_IN + i] * L.f[r][i];
    const float sg = sigmoidf(z), h = z * sg;
    L.h1[r][n] = h;
    L.d1[r][n] = sg + h * (1.0f - sg);
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = prm[CV_B2 + n];
    for (int k = 0; k < CV_H; ++k) z += prm[CV_W2 + n * CV_H + k] * L.h1[r][k];
    const float sg = sigmoidf(z), h = z * sg;
    L.h2[r][n] = h;
    L.d2[r][n] = sg + h * (1.0f - sg);
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = prm[CV_B3 + n];
    for (int k = 0; k < CV_H; ++k) z += prm[CV_W3 + n * CV_H + k] * L.h2[r][k];
    const float sg = sigmoidf(z), h = z * sg;
    L.h3[r][n] = h;
    L.d3[r][n] = sg + h * (1.0f - sg);
  }
  __syncthreads();
  // head, targets and output errors
  float lsum = 0.0f;
  for (int a = 0; a < 8; ++a) {
    const int k = q8 * 8 + a;
    float y = prm[CV_B4 + k];
    for (int n = 0; n < CV_H; ++n) y += prm[CV_W4 + k * CV_H + n] * L.h3[r][n];
    float tg = 0.0f;
    bool live;
    if (fixed_f) {
      live = k < nk_all;
      tg = live ? fixed_t[row * CV_OUT + k] : 0.0f;
    } else {
      const CvSet &S = sets[L.set[r]];
      live = k < S.nk;
      if (live) {
        const double pv = cv_target(S, L.path[r], k) / S.disc_of[L.cp[r]];
        tg = static_cast<float>((pv - static_cast<double>(S.mu[k])) * static_cast<double>(S.inv_sd[k]));
      }
    }
    const float e = live ? y - tg : 0.0f;
    lsum += e * e;
    L.dy[r][k] = 2.0f * e * inv_n;
  }
  L.loss[r][q8] = lsum;
  __syncthreads();
  // backward: deltas overwrite the activation derivatives
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float s = 0.f;
    for (int k = 0; k < CV_OUT; ++k) s += prm[CV_W4 + k * CV_H + n] * L.dy[r][k];
    L.d3[r][n] *= s;
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int k = q8 * 8 + a;
    float s = 0.f;
    for (int n = 0; n < CV_H; ++n) s += prm[CV_W3 + n * CV_H + k] * L.d3[r][n];
    L.d2[r][k] *= s;
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int k = q8 * 8 + a;
    float s = 0.f;
    for (int n = 0; n < CV_H; ++n) s += prm[CV_W2 + n * CV_H + k] * L.d2[r][n];
    L.d1[r][k] *= s;
  }
  __syncthreads();
  // parameter gradients of this block's rows
  float *g = gpart + static_cast<size_t>(blockIdx.x) * CV_NPARAM;
  for (int e = tid; e < CV_H * CV_H; e += BLOCK) {
    const int n = e / CV_H, k = e % CV_H;
    float s2 = 0.f, s3 = 0.f, s4 = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr) {
      s2 += L.d2[rr][n] * L.h1[rr][k];
      s3 += L.d3[rr][n] * L.h2[rr][k];
      s4 += L.dy[rr][n] * L.h3[rr][k];
    }
    g[CV_W2 + e] = s2;
    g[CV_W3 + e] = s3;
    g[CV_W4 + e] = s4;
  }
  for (int e = tid; e < CV_H * CV_IN; e += BLOCK) {
    const int n = e / CV_IN, i = e % CV_IN;
    float s = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr) s += L.d1[rr][n] * L.f[rr][i];
    g[CV_W1 + e] = s;
  }
  for (int n = tid; n < CV_H; n += BLOCK) {
    float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr) {
      s1 += L.d1[rr][n];
      s2 += L.d2[rr][n];
      s3 += L.d3[rr][n];
      s4 += L.dy[rr][n];
    }
    g[CV_B1 + n] = s1;
    g[CV_B2 + n] = s2;
    g[CV_B3 + n] = s3;
    g[CV_B4 + n] = s4;
  }
  if (tid == 0) {
    float l = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr)
      for (int q = 0; q < 8; ++q) l += L.loss[rr][q];
    lpart[blockIdx.x] = l;
  }
}

// Adam on the block-summed gradient (fixed block order). state: m[NPARAM],
// v[NPARAM]; step counter read here, advanced by cv_step_kernel.
__global__ void cv_adam_kernel(float *__restrict__ prm, float *__restrict__ m, float *__restrict__ v,
                               const float *__restrict__ gpart, int n_blocks, const int *__restrict__ step, float lr,
                               const float *__restrict__ lpart, float *__restrict__ loss_hist, float loss_scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = *step + 1;
  if (i < CV_NPARAM) {
    float g = 0.f;
    for (int b = 0; b < n_blocks; ++b) g += gpart[static_cast<size_t>(b) * CV_NPARAM + i];
    const float b1 = 0.9f, b2 = 0.999f;
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, static_cast<float>(t)));
    const float vh = vi / (1.f - powf(b2, static_cast<float>(t)));
    prm[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
  }
  if (i == 0) {
    float l = 0.f;
    for (int b = 0; b < n_blocks; ++b) l += lpart[b];
    loss_hist[t - 1] = l * loss_scale;
  }
}

__global__ void cv_step_kernel(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// debug: FP16 MFMA tangents of n rows of features (rows beyond n padded
// with zeros): out[(row * 64 + k) * 3 + t] = d out_k / d f_{1+t}
__global__ void __launch_bounds__(BLOCK) cv_grad_debug_kernel(const float *__restrict__ prm,
                                                              const float *__restrict__ feats, int n,
                                                              float *__restrict__ out) {
  __shared__ CvLds L;
  cv_load_lds(L, prm, nullptr, 1.0f, CV_OUT);
  __syncthreads();
  const int l = threadIdx.x & 63;
  const int wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64;
  const int row = wave * 16 + (l & 15);
  float f[CV_IN];
  for (int i = 0; i < CV_IN; ++i) f[i] = row < n ? feats[row * CV_IN + i] : 0.f;
  cv_h4 dd[3][4];
  cv_value(L, f, dd, l);
  for (int t = 0; t < 3; ++t) {
    cv_v4f tg[4], g[4];
    cv_tangent(L, dd, 1 + t, tg, l);
    cv_head<false>(L.hd[0], tg, g, l);
    if (row < n)
      for (int m = 0; m < 4; ++m)
        for (int e = 0; e < 4; ++e) out[(static_cast<size_t>(row) * CV_OUT + cv_unit(m, e, l)) * 3 + t] = g[m][e];
  }
}

}  // namespace myql
