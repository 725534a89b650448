// ORACLE (test infrastructure only): CPU FP64 restatement of the reference's
// path simulation -- variance schemes, jumps, the ASVJ stepper and its MLMC
// coupled twin, path trackers and payoffs.
//
//   SchemeNV     VolSchemes.hpp:24-66
//   SchemeExact  VolSchemes.hpp:71-225 (Marsaglia-Tsang for alpha > 1,
//                Ahrens-Dieter GS for alpha <= 1)
//   SchemeNCI    VolSchemes.hpp:230-377 (chi-squared quantile table with
//                Fritsch-Butland slopes and cubic Hermite interpolation)
//   MertonJump / KouJump   JumpPolicies.hpp:32-135
//   ASVJStepper  ASVJstepper.hpp:85-229, init_cir_impl :262-283
//   CoupledASVJStepper::multi_step_coupled  CoupledASVJStepper.hpp:170-517
//   Trackers     PathTrackers.hpp:8-300
//   Payoffs      Payoffs.hpp:14-69
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/myql_capi.h"
#include "orc_numerics.hpp"

namespace orc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------------------
// NCI global workspace: chi-squared quantile rows y[k][i] for dof d + 2k at
// u = i * dx, with harmonic-mean (Fritsch-Butland) slopes m[k][i].
// ---------------------------------------------------------------------------
struct NciTable {
  double d = 0.0;
  size_t n_max = 0, grid = 10000;
  double dx = 0.0, inv_dx = 0.0;
  std::vector<double> y, m;
  // Upper-tail rows (u >= 0.999) on w = -log(1 - u): quantiles at
  // w_j = w0 + j hw and exact slopes dq/dw = (1 - u) / pdf(q). Empty under
  // the reference's table semantics (exact inversion there).
  static constexpr size_t kTail = 512;
  static constexpr double kTailW1 = 38.0;
  std::vector<double> ty, tm;
  double tw0 = 0.0, thw = 0.0, inv_thw = 0.0;
};

// Row count: the reference's rule (VolSchemes.hpp:245-255, v_max = max(v0,
// 3 theta), margin 4 sqrt(l)) or the engine's default extended rule (v =
// max(v0, 25 theta), margin 6 sqrt(l)); both clamped to [23, 1023].
static inline size_t nci_rows_nmax(const myql_heston_params &p, double dt, bool reference_rows) {
  const double v_max = std::max(p.v0, (reference_rows ? 3.0 : 25.0) * p.theta);
  const double ex = std::exp(-p.kappa * dt);
  const double psi = (std::fabs(p.kappa) < 1e-10) ? dt : (1.0 - ex) / p.kappa;
  const double lam_max = 2.0 * v_max * ex / (p.sigma * p.sigma * psi);
  size_t n_max = static_cast<size_t>(std::ceil(lam_max + (reference_rows ? 4.0 : 6.0) * std::sqrt(lam_max)));
  if (n_max < 23) n_max = 23;
  if (n_max > 1023) n_max = 1023;
  return n_max;
}

// Rows up to n_fb get the reference's Fritsch-Butland slopes; rows above it
// (present only outside the reference's table semantics) get the exact
// slopes dq/du = 1/pdf(q).
static inline std::shared_ptr<NciTable> build_nci_table(const myql_heston_params &p, double dt,
                                                        bool reference_rows, long n_fb = -1) {
  const OpsPause setup_work;
  auto gw = std::make_shared<NciTable>();
  gw->d = 4.0 * p.kappa * p.theta / (p.sigma * p.sigma);
  double nm = static_cast<double>(nci_rows_nmax(p, dt, reference_rows));
  size_t n_max = static_cast<size_t>(nm);
  if (n_max < 23) n_max = 23;
  if (n_max > 1023) n_max = 1023;
  gw->n_max = n_max;
  gw->grid = 10000;
  gw->dx = 0.999 / static_cast<double>(gw->grid - 1);
  gw->inv_dx = 1.0 / gw->dx;
  const size_t G = gw->grid;
  gw->y.assign((n_max + 1) * G, 0.0);
  gw->m.assign((n_max + 1) * G, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (long k = 0; k <= static_cast<long>(n_max); ++k) {
    const double dof = gw->d + 2.0 * static_cast<double>(k);
    double *y = &gw->y[k * G];
    double *m = &gw->m[k * G];
    for (size_t i = 0; i < G; ++i) y[i] = chi2_quantile(dof, static_cast<double>(i) * gw->dx);
    const bool exact_slopes = !reference_rows && n_fb >= 0 && k > n_fb;
    const double lgh = std::lgamma(0.5 * dof);
    for (size_t i = 0; i < G; ++i) {
      if (exact_slopes && i > 0) {
        const double lpdf = (0.5 * dof - 1.0) * std::log(y[i]) - 0.5 * y[i] - 0.5 * dof * std::log(2.0) - lgh;
        m[i] = 1.0 / std::exp(lpdf);
      } else if (i == 0) {
        m[i] = (y[1] - y[0]) * gw->inv_dx;
      } else if (i == G - 1) {
        m[i] = (y[i] - y[i - 1]) * gw->inv_dx;
      } else {
        const double d1 = (y[i] - y[i - 1]) * gw->inv_dx;
        const double d2 = (y[i + 1] - y[i]) * gw->inv_dx;
        m[i] = (d1 * d2 > 0.0) ? 2.0 * d1 * d2 / (d1 + d2) : 0.0;
      }
    }
  }
  if (!reference_rows) {
    const size_t P = NciTable::kTail;
    gw->tw0 = std::log(1000.0);
    gw->thw = (NciTable::kTailW1 - gw->tw0) / static_cast<double>(P - 1);
    gw->inv_thw = 1.0 / gw->thw;
    gw->ty.assign((n_max + 1) * P, 0.0);
    gw->tm.assign((n_max + 1) * P, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (long k = 0; k <= static_cast<long>(n_max); ++k) {
      const double dof = gw->d + 2.0 * static_cast<double>(k);
      const double lgh = std::lgamma(0.5 * dof);
      for (size_t j = 0; j < P; ++j) {
        const double w = gw->tw0 + static_cast<double>(j) * gw->thw;
        const double qt = std::exp(-w);
        const double x = chi2_quantile_upper(dof, qt);
        const double lpdf = (0.5 * dof - 1.0) * std::log(x) - 0.5 * x - 0.5 * dof * std::log(2.0) - lgh;
        gw->ty[k * P + j] = x;
        gw->tm[k * P + j] = qt / std::exp(lpdf);
      }
    }
  }
  return gw;
}

// ---------------------------------------------------------------------------
// One CIR factor prepared for a step size: log-S coefficients C[0..3] and the
// scheme's constants.
// ---------------------------------------------------------------------------
struct Factor {
  int scheme = MYQL_SCHEME_NULL;
  myql_heston_params p{};
  double C[4] = {0, 0, 0, 0};  // logS += C0 + C1 v + C2 v' + sqrt(C3 (v+v')) G
  double K[3] = {0, 0, 0};     // scheme constants
  std::shared_ptr<NciTable> nci;
  // SchemeExact small-alpha (GS) constants
  bool alpha_small = false;
  double sm_p = 0, sm_inv_p = 0, sm_1mp = 0, sm_1ma = 0, sm_inv_a = 0;
};

static inline void prepare_factor(Factor &f, int scheme, const myql_heston_params &p, double r_part,
                                  double h, std::shared_ptr<NciTable> nci) {
  f.scheme = scheme;
  f.p = p;
  const double kappa = p.kappa, theta = p.theta, sigma = p.sigma, rho = p.rho;
  f.C[0] = (r_part - rho * kappa * theta / sigma) * h;
  const double term = (kappa * rho / sigma - 0.5) * h * 0.5;
  f.C[1] = term - rho / sigma;
  f.C[2] = term + rho / sigma;
  f.C[3] = h * (1.0 - rho * rho) * 0.5;
  if (scheme == MYQL_SCHEME_NV) {
    if (sigma * sigma > 4.0 * kappa * theta)
      throw Error(MYQL_E_NV_FELLER, "SchemeNV violation: sigma^2 > 4*kappa*theta");
    const double ex = std::exp(-kappa * h * 0.5);
    const double psi = (std::fabs(kappa) < 1e-10) ? h * 0.5 : (1.0 - ex) / kappa;
    const double D = kappa * theta - sigma * sigma * 0.25;
    f.K[0] = ex;
    f.K[1] = D * psi;
    f.K[2] = sigma * std::sqrt(h) * 0.5;
  } else if (scheme == MYQL_SCHEME_NCI || scheme == MYQL_SCHEME_EXACT) {
    const double ex = std::exp(-kappa * h);
    const double psi = (std::fabs(kappa) < 1e-10) ? h : (1.0 - ex) / kappa;
    f.K[0] = (2.0 / (sigma * sigma * psi)) * ex;
    if (scheme == MYQL_SCHEME_NCI) {
      f.K[1] = 4.0 * kappa * theta / (sigma * sigma);
      f.K[2] = 0.25 * sigma * sigma * psi;
      f.nci = nci;
    } else {
      const double ab = 2.0 * kappa * theta / (sigma * sigma);
      f.K[1] = ab;
      f.K[2] = 0.5 * sigma * sigma * psi;
      f.alpha_small = ab <= 1.0;
      if (f.alpha_small) {
        const double e = 2.71828182845904523536;
        f.sm_p = e / (ab + e);
        f.sm_inv_p = 1.0 / f.sm_p;
        f.sm_1mp = 1.0 - f.sm_p;
        f.sm_1ma = 1.0 - ab;
        f.sm_inv_a = 1.0 / ab;
      }
    }
  }
}

static inline double nv_coupled(const Factor &f, double v, double Z) {
  const double root = std::sqrt(f.K[0] * v + f.K[1]) + f.K[2] * Z;
  ORC_OPS(8);  // FMA, sqrt, FMA, square, FMA
  return f.K[0] * (root * root) + f.K[1];
}

static inline double exact_gamma(const Factor &f, int N, Xoshiro &g) {
  const double alpha = f.K[1] + N;
  if (alpha > 1.0) {
    const double d = alpha - 1.0 / 3.0;
    const double c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double Z, v;
      do {
        Z = normcdfinv(uniform(g));
        v = 1.0 + c * Z;
        ORC_OPS(2);
      } while (v <= 0.0);
      v = v * v * v;
      const double U = uniform(g);
      ORC_OPS(7);  // cube, squeeze test
      if (U < 1.0 - 0.0331 * Z * Z * Z * Z) return d * v;
      ORC_OPS(8);  // log test
      if (std::log(U) < 0.5 * Z * Z + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
  for (;;) {
    const double u = uniform(g);
    const double y = exponential(1.0, g);
    ORC_OPS(6);
    if (u < f.sm_p) {
      const double x = std::exp(-y * f.sm_inv_a);
      if (u * f.sm_inv_p < 1.0 - x) return x;
      if (u < f.sm_p * std::exp(-x)) return x;
    } else {
      const double x = 1.0 + y;
      const double inv_h = 1.0 / (1.0 + f.sm_1ma * y);
      if (u < f.sm_p + f.sm_1mp * inv_h) return x;
      if (u < f.sm_p + f.sm_1mp * std::pow(x, -f.sm_1ma)) return x;
    }
  }
}

// diagnostics: how often the exact quantile replaces the table
inline std::atomic<uint64_t> &nci_fallbacks() {
  static std::atomic<uint64_t> n{0};
  return n;
}

// Outside the reference's table semantics (tail rows present), rows N >= 2
// are inverted exactly in the first grid interval U < dx: their quantile is 0
// at the u = 0 node and rises like u^(2/(d+2N)) just above it, so the
// interpolated first interval biases E[v'] low (by up to ~3e-5 of the row
// mean at large N), an error each step adds to the mean-reverting variance.
static inline double nci_lookup(const Factor &f, int N, double U) {
  const NciTable &t = *f.nci;
  const bool lo_exact = !t.ty.empty() && N >= 2 && U < t.dx;
  if (static_cast<size_t>(N) <= t.n_max && U < 0.999 && !lo_exact) {
    const double pos = U * t.inv_dx;
    size_t i = static_cast<size_t>(pos);
    if (i >= t.grid - 1) i = t.grid - 2;
    const double tt = pos - static_cast<double>(i);
    const size_t off = static_cast<size_t>(N) * t.grid + i;
    const double y0 = t.y[off], y1 = t.y[off + 1], m0 = t.m[off], m1 = t.m[off + 1];
    const double t2 = tt * tt, t3 = t2 * tt;
    ORC_OPS(28);  // position, Hermite basis, 4-term combination, scale
    return ((2 * t3 - 3 * t2 + 1) * y0 + (t3 - 2 * t2 + tt) * t.dx * m0 + (-2 * t3 + 3 * t2) * y1 +
            (t3 - t2) * t.dx * m1) *
           f.K[2];
  }
  if (static_cast<size_t>(N) <= t.n_max && !t.ty.empty() && U >= 0.999) {
    // upper tail: cubic Hermite in w = -log(1 - U) with exact slopes
    const size_t P = NciTable::kTail;
    double pos = (-std::log1p(-U) - t.tw0) * t.inv_thw;
    if (pos < 0.0) pos = 0.0;
    size_t j = static_cast<size_t>(pos);
    if (j >= P - 1) j = P - 2;
    const double tt = pos - static_cast<double>(j);
    const size_t off = static_cast<size_t>(N) * P + j;
    const double t2 = tt * tt, t3 = t2 * tt;
    ORC_OPS(31);  // as above plus the log1p of the tail coordinate
    return ((2 * t3 - 3 * t2 + 1) * t.ty[off] + (t3 - 2 * t2 + tt) * t.thw * t.tm[off] +
            (-2 * t3 + 3 * t2) * t.ty[off + 1] + (t3 - t2) * t.thw * t.tm[off + 1]) *
           f.K[2];
  }
  nci_fallbacks().fetch_add(1, std::memory_order_relaxed);
  ORC_OPS(3);
  return chi2_quantile(f.K[1] + 2.0 * N, U) * f.K[2];
}

// Scheme::evolve (the draws it consumes are the scheme's own).
static inline double evolve(const Factor &f, double v, Xoshiro &g) {
  switch (f.scheme) {
    case MYQL_SCHEME_NV: {
      const double Z = normcdfinv(uniform(g));
      return nv_coupled(f, v, Z);
    }
    case MYQL_SCHEME_NCI: {
      ORC_OPS(1);  // Poisson intensity
      const int N = static_cast<int>(poissinvcdf(f.K[0] * v, uniform(g)));
      const double U = uniform(g);
      return nci_lookup(f, N, U);
    }
    case MYQL_SCHEME_EXACT: {
      ORC_OPS(2);  // intensity, scale
      const int N = static_cast<int>(poissinvcdf(f.K[0] * v, uniform(g)));
      return exact_gamma(f, N, g) * f.K[2];
    }
  }
  return v;
}

// ---------------------------------------------------------------------------
// Jumps: workspace prepared for a time span (dt for path-dependent trackers,
// the whole maturity for Europeans).
// ---------------------------------------------------------------------------
struct JumpWs {
  int kind = MYQL_JUMP_NONE;
  double mean = 0, exp_mean = 1, drift = 0;
  double a = 0, b = 0, c = 0, d = 0;  // Merton mu, delta; Kou p, eta1, eta2
};

static inline JumpWs prepare_jump(const myql_model &m, double span) {
  JumpWs w;
  w.kind = m.jump_kind;
  if (m.jump_kind == MYQL_JUMP_MERTON) {
    const double lam = m.jump[0], mu = m.jump[1], delta = m.jump[2];
    w.mean = lam * span;
    w.exp_mean = std::exp(-w.mean);
    const double k = std::exp(mu + 0.5 * delta * delta) - 1.0;
    w.drift = -lam * k * span;
    w.a = mu;
    w.b = delta;
  } else if (m.jump_kind == MYQL_JUMP_KOU) {
    const double lam = m.jump[0], p = m.jump[1], e1 = m.jump[2], e2 = m.jump[3];
    w.mean = lam * span;
    w.exp_mean = std::exp(-w.mean);
    const double k = (p * e1) / (e1 - 1.0) + ((1.0 - p) * e2) / (e2 + 1.0) - 1.0;
    w.drift = -lam * k * span;
    w.a = p;
    w.b = e1;
    w.c = e2;
  }
  return w;
}

static inline double log_jump(const JumpWs &w, Xoshiro &g) {
  if (w.kind == MYQL_JUMP_NONE) return 0.0;
  const int n = jump_count(w.mean, w.exp_mean, g);
  if (n == 0) return w.drift;
  if (w.kind == MYQL_JUMP_MERTON) {
    const double Z = normcdfinv(uniform(g));
    ORC_OPS(6);
    return w.drift + n * w.a + std::sqrt(static_cast<double>(n)) * w.b * Z;
  }
  double total = 0.0;
  for (int i = 0; i < n; ++i) {
    const double u = uniform(g);
    if (u < w.a) total += exponential(w.b, g);
    else total -= exponential(w.c, g);
  }
  ORC_OPS(n + 1);
  return w.drift + total;
}

// ---------------------------------------------------------------------------
// Trackers
// ---------------------------------------------------------------------------
struct TrackerCfg {
  int kind = MYQL_INSTR_EUROPEAN;
  bool up = true;  // barrier direction
  double past_avg_log = 0, past_avg_S = 0, t_elapsed = 0;
  double hist_min_logS = 1e300, hist_max_logS = -1e300;
  double b_base = 0, b_up = 0, b_dn = 0;
};

struct PathState {
  double logS = 0;  // log price; the price itself after finalize
  double V[2] = {0, 0};
  double sumLogS = 0, sumS = 0, minLogS = 0, maxLogS = 0;
  double t = 0;
  bool hit[3] = {false, false, false};  // base, up, dn
  double hit_t[3] = {0, 0, 0};
};

static inline bool path_dependent(int kind) { return kind != MYQL_INSTR_EUROPEAN; }

static inline void trk_init(PathState &s, const TrackerCfg &c, double logS0, double dt) {
  switch (c.kind) {
    case MYQL_INSTR_GEO_ASIAN:
      s.sumLogS = c.past_avg_log * c.t_elapsed + 0.5 * s.logS * dt;
      break;
    case MYQL_INSTR_ARITH_ASIAN:
      s.sumLogS = c.past_avg_log * c.t_elapsed + 0.5 * s.logS * dt;
      s.sumS = c.past_avg_S * c.t_elapsed + 0.5 * std::exp(s.logS) * dt;
      break;
    case MYQL_INSTR_LOOKBACK:
      s.minLogS = (c.hist_min_logS < 1e300) ? c.hist_min_logS : logS0;
      s.maxLogS = (c.hist_max_logS > -1e300) ? c.hist_max_logS : logS0;
      break;
    case MYQL_INSTR_BARRIER: {
      s.t = 0.0;
      const double lv[3] = {c.b_base, c.b_up, c.b_dn};
      for (int i = 0; i < 3; ++i) {
        s.hit[i] = c.up ? (logS0 >= lv[i]) : (logS0 <= lv[i]);
        s.hit_t[i] = 0.0;
      }
      break;
    }
    default:
      break;
  }
}

static inline void mark_hit(PathState &s, int i) {
  if (!s.hit[i]) {
    s.hit[i] = true;
    s.hit_t[i] = s.t;
  }
}

static inline void trk_update(PathState &s, const TrackerCfg &c, double dt) {
  switch (c.kind) {
    case MYQL_INSTR_GEO_ASIAN:
      s.sumLogS += s.logS * dt;
      ORC_OPS(2);
      break;
    case MYQL_INSTR_ARITH_ASIAN:
      s.sumLogS += s.logS * dt;
      s.sumS += std::exp(s.logS) * dt;
      ORC_OPS(5);
      break;
    case MYQL_INSTR_LOOKBACK:
      s.maxLogS = (s.logS > s.maxLogS) ? s.logS : s.maxLogS;
      s.minLogS = (s.logS < s.minLogS) ? s.logS : s.minLogS;
      break;
    case MYQL_INSTR_BARRIER:
      s.t += dt;
      ORC_OPS(1);
      // nested tests, easiest barrier first (PathTrackers.hpp:250-294)
      if (c.up) {
        if (s.logS >= c.b_up) {
          mark_hit(s, 1);
          if (s.logS >= c.b_base) {
            mark_hit(s, 0);
            if (s.logS >= c.b_dn) mark_hit(s, 2);
          }
        }
      } else {
        if (s.logS <= c.b_dn) {
          mark_hit(s, 2);
          if (s.logS <= c.b_base) {
            mark_hit(s, 0);
            if (s.logS <= c.b_up) mark_hit(s, 1);
          }
        }
      }
      break;
    default:
      break;
  }
}

static inline void trk_finalize(PathState &s, const TrackerCfg &c, double dt, double T) {
  switch (c.kind) {
    case MYQL_INSTR_GEO_ASIAN: {
      s.sumLogS -= 0.5 * s.logS * dt;
      s.sumLogS = std::exp(s.sumLogS / (T + c.t_elapsed));
      break;
    }
    case MYQL_INSTR_ARITH_ASIAN: {
      s.sumLogS -= 0.5 * s.logS * dt;
      s.sumS -= 0.5 * std::exp(s.logS) * dt;
      const double Tt = T + c.t_elapsed;
      s.sumLogS = std::exp(s.sumLogS / Tt);
      s.sumS = s.sumS / Tt;
      break;
    }
    case MYQL_INSTR_LOOKBACK:
      s.minLogS = std::exp(s.minLogS);
      s.maxLogS = std::exp(s.maxLogS);
      break;
    default:
      break;
  }
  s.logS = std::exp(s.logS);
  ORC_OPS(c.kind == MYQL_INSTR_EUROPEAN ? 1 : 7);
}

// ---------------------------------------------------------------------------
// Stepper prepared for (model, step dt, r, q, total time, workspaces).
// ---------------------------------------------------------------------------
struct Stepper {
  int nf = 0;
  double dt = 0;
  double gbm_drift = 0, gbm_vol = 0;
  double v0[2] = {0, 0};
  Factor f[2];
  JumpWs jump;
  bool pathdep = false;
};

static inline Stepper make_stepper(const myql_model &m, const int32_t schemes[2], double dt, double r,
                                   double q, double total_time, bool pathdep,
                                   std::shared_ptr<NciTable> nci0, std::shared_ptr<NciTable> nci1) {
  Stepper st;
  st.nf = m.n_factors;
  st.dt = dt;
  st.pathdep = pathdep;
  st.jump = prepare_jump(m, pathdep ? dt : total_time);
  if (st.nf == 0) {
    const double var = m.vol * m.vol;
    st.gbm_drift = (r - q - 0.5 * var) * dt;
    st.gbm_vol = m.vol * std::sqrt(dt);
  } else {
    const double r_part = (r - q) / static_cast<double>(st.nf);
    for (int i = 0; i < st.nf; ++i) {
      prepare_factor(st.f[i], schemes[i], m.heston[i], r_part, dt, i == 0 ? nci0 : nci1);
      st.v0[i] = m.heston[i].v0;
    }
  }
  return st;
}

static inline void start_path(const Stepper &st, PathState &s, const TrackerCfg &c, double S0) {
  s = PathState();
  s.logS = std::log(S0);
  s.V[0] = st.v0[0];
  s.V[1] = st.v0[1];
  trk_init(s, c, s.logS, st.dt);
}

// ASVJStepper::multi_step. rec (debug exports only): per step, the index of
// its first uniform in the draw log, then log S, V0, V1 and the step's log
// jump after the step (PLAIN_REC doubles); the European's terminal jump in
// one more row.
constexpr int PLAIN_REC = 5;
static inline void multi_step(const Stepper &st, PathState &s, const TrackerCfg &c, Xoshiro &g,
                              unsigned n, double *rec = nullptr) {
  double v0 = s.V[0], v1 = s.V[1];
  double x = s.logS;
  const double total_time = n * st.dt;
  if (st.nf == 0 && !st.pathdep) {
    const double G = normcdfinv(uniform(g));
    x += st.gbm_drift * n + st.gbm_vol * std::sqrt(static_cast<double>(n)) * G;
    ORC_OPS(6);
    x += log_jump(st.jump, g);
    s.logS = x;
    trk_finalize(s, c, st.dt, total_time);
    return;
  }
  for (unsigned k = 0; k < n; ++k) {
    double *rk = rec ? rec + static_cast<size_t>(k) * PLAIN_REC : nullptr;
    if (rk) rk[0] = draw_index();
    double jk = 0.0;
    if (st.nf == 0) {
      const double G = normcdfinv(uniform(g));
      x += st.gbm_drift + st.gbm_vol * G;
      ORC_OPS(3);
    } else if (st.nf == 1) {
      const double *C = st.f[0].C;
      const double G = normcdfinv(uniform(g));
      const double vn = evolve(st.f[0], v0, g);
      const double root = std::sqrt(C[3] * (v0 + vn));
      x += C[0] + C[1] * v0 + C[2] * vn + root * G;
      ORC_OPS(10);  // log-S step: C3 (v + v'), sqrt, three FMAs, adds
      v0 = vn;
    } else {
      const double *C0 = st.f[0].C, *C1 = st.f[1].C;
      const double G = normcdfinv(uniform(g));
      const double vn0 = evolve(st.f[0], v0, g);
      const double vn1 = evolve(st.f[1], v1, g);
      const double root = std::sqrt(C0[3] * (v0 + vn0) + C1[3] * (v1 + vn1));
      x += C0[0] + C0[1] * v0 + C0[2] * vn0 + C1[0] + C1[1] * v1 + C1[2] * vn1 + root * G;
      ORC_OPS(19);
      v0 = vn0;
      v1 = vn1;
    }
    if (st.pathdep) {
      jk = log_jump(st.jump, g);
      x += jk;
      s.logS = x;
      trk_update(s, c, st.dt);
    }
    if (rk) {
      rk[1] = x;
      rk[2] = v0;
      rk[3] = v1;
      rk[4] = jk;
    }
  }
  if (!st.pathdep && st.nf > 0) {
    const double J = log_jump(st.jump, g);
    x += J;
    if (rec) rec[static_cast<size_t>(n) * PLAIN_REC] = J;
  }
  s.logS = x;
  s.V[0] = v0;
  s.V[1] = v1;
  trk_finalize(s, c, st.dt, total_time);
}

// CoupledASVJStepper::multi_step_coupled. `fine` is prepared with dt_f,
// `coarse` with 2 dt_f; their jump workspaces are the fine one's. rec
// (debug exports only): per coarse step, after it, COUPLED_REC doubles:
// draw-log index of its first uniform, G1, G2, Gc, a1, a2, j1, j2, xf, vf0,
// vf1, xc, vc0, vc1, mid0, mid1 (mid = fine variance after the first fine
// step); the European's shared jump in one more row.
constexpr int COUPLED_REC = 16;
static inline void multi_step_coupled(const Stepper &F, const Stepper &Cs, PathState &sf, PathState &sc,
                                      const TrackerCfg &c, Xoshiro &g, unsigned nc, double *rec = nullptr) {
  double vf[2] = {sf.V[0], sf.V[1]}, vc[2] = {sc.V[0], sc.V[1]};
  double xf = sf.logS, xc = sc.logS;
  const double total_time = nc * Cs.dt;
  const bool pd = F.pathdep;
  const double SQ2 = 1.4142135623730951;
  if (F.nf == 0 && !pd) {
    const double G = normcdfinv(uniform(g));
    xf += F.gbm_drift * (nc * 2) + F.gbm_vol * std::sqrt(static_cast<double>(nc * 2)) * G;
    xc += Cs.gbm_drift * nc + Cs.gbm_vol * std::sqrt(static_cast<double>(nc)) * G;
    const double J = log_jump(F.jump, g);
    xf += J;
    xc += J;
    sf.logS = xf;
    sc.logS = xc;
    trk_finalize(sf, c, F.dt, total_time);
    trk_finalize(sc, c, Cs.dt, total_time);
    return;
  }
  if (!pd && F.nf > 0) {
    const double J = log_jump(F.jump, g);
    xf += J;
    xc += J;
    if (rec) rec[static_cast<size_t>(nc) * COUPLED_REC] = J;
  }
  for (unsigned k = 0; k < nc; ++k) {
    double *rk = rec ? rec + static_cast<size_t>(k) * COUPLED_REC : nullptr;
    if (rk) rk[0] = draw_index();
    const double G1 = normcdfinv(uniform(g));
    const double G2 = normcdfinv(uniform(g));
    double Gc = (G1 + G2) / SQ2;
    double j1 = 0.0, j2 = 0.0;
    double ra1 = 0.0, ra2 = 0.0, rmid[2] = {0.0, 0.0};
    if (F.nf == 0) {
      xc += Cs.gbm_drift + Cs.gbm_vol * Gc;
      xf += F.gbm_drift + F.gbm_vol * G1;
      if (pd) {
        j1 = log_jump(F.jump, g);
        xf += j1;
        sf.logS = xf;
        trk_update(sf, c, F.dt);
      }
      xf += F.gbm_drift + F.gbm_vol * G2;
      if (pd) {
        j2 = log_jump(F.jump, g);
        xf += j2;
        sf.logS = xf;
        trk_update(sf, c, F.dt);
        xc += j1 + j2;
        sc.logS = xc;
        trk_update(sc, c, Cs.dt);
      }
    } else {
      const int nf = F.nf;
      double mid[2], nxt[2], Z1[2] = {0, 0}, Z2[2] = {0, 0};
      // fine step 1
      for (int i = 0; i < nf; ++i) {
        if (F.f[i].scheme == MYQL_SCHEME_NV) {
          Z1[i] = normcdfinv(uniform(g));
          mid[i] = nv_coupled(F.f[i], vf[i], Z1[i]);
        } else {
          mid[i] = evolve(F.f[i], vf[i], g);
        }
      }
      double a1sq = 0.0, drift1 = 0.0;
      for (int i = 0; i < nf; ++i) {
        const double *C = F.f[i].C;
        a1sq += C[3] * (vf[i] + mid[i]);
        drift1 += C[0] + C[1] * vf[i] + C[2] * mid[i];
      }
      const double a1 = std::sqrt(a1sq);
      xf += drift1 + a1 * G1;
      if (pd) {
        j1 = log_jump(F.jump, g);
        xf += j1;
        sf.logS = xf;
        trk_update(sf, c, F.dt);
      }
      // fine step 2
      for (int i = 0; i < nf; ++i) {
        if (F.f[i].scheme == MYQL_SCHEME_NV) {
          Z2[i] = normcdfinv(uniform(g));
          nxt[i] = nv_coupled(F.f[i], mid[i], Z2[i]);
        } else {
          nxt[i] = evolve(F.f[i], mid[i], g);
        }
      }
      double a2sq = 0.0, drift2 = 0.0;
      for (int i = 0; i < nf; ++i) {
        const double *C = F.f[i].C;
        a2sq += C[3] * (mid[i] + nxt[i]);
        drift2 += C[0] + C[1] * mid[i] + C[2] * nxt[i];
      }
      const double a2 = std::sqrt(a2sq);
      xf += drift2 + a2 * G2;
      if (pd) {
        j2 = log_jump(F.jump, g);
        xf += j2;
        sf.logS = xf;
        trk_update(sf, c, F.dt);
      }
      // coarse step
      double cn[2];
      for (int i = 0; i < nf; ++i) {
        if (F.f[i].scheme == MYQL_SCHEME_NV)
          cn[i] = nv_coupled(Cs.f[i], vc[i], (Z1[i] + Z2[i]) / SQ2);
        else
          cn[i] = nxt[i];
      }
      if (a1 + a2 > 1e-15) Gc = (a1 * G1 + a2 * G2) / std::hypot(a1, a2);
      double acsq = 0.0, driftc = 0.0;
      for (int i = 0; i < nf; ++i) {
        const double *C = Cs.f[i].C;
        acsq += C[3] * (vc[i] + cn[i]);
        driftc += C[0] + C[1] * vc[i] + C[2] * cn[i];
      }
      xc += driftc + std::sqrt(acsq) * Gc;
      for (int i = 0; i < nf; ++i) {
        vf[i] = nxt[i];
        vc[i] = cn[i];
      }
      if (pd) {
        xc += j1 + j2;
        sc.logS = xc;
        trk_update(sc, c, Cs.dt);
      }
      ra1 = a1;
      ra2 = a2;
      rmid[0] = mid[0];
      rmid[1] = nf > 1 ? mid[1] : 0.0;
    }
    if (rk) {
      const double r[COUPLED_REC - 1] = {G1, G2, Gc, ra1, ra2, j1, j2, xf, vf[0], vf[1], xc, vc[0], vc[1],
                                          rmid[0], rmid[1]};
      for (int i = 0; i < COUPLED_REC - 1; ++i) rk[1 + i] = r[i];
    }
  }
  sf.logS = xf;
  sf.V[0] = vf[0];
  sf.V[1] = vf[1];
  sc.logS = xc;
  sc.V[0] = vc[0];
  sc.V[1] = vc[1];
  trk_finalize(sf, c, F.dt, total_time);
  trk_finalize(sc, c, Cs.dt, total_time);
}

// ---------------------------------------------------------------------------
// Payoffs and instrument buffers
// ---------------------------------------------------------------------------
struct Instr {
  myql_instrument in;
  std::vector<double> K;
  double eps = -1.0;
};

static inline double payoff(const Instr &I, int mode, double spot, double strike) {
  const double diff = (I.in.option_type == MYQL_CALL) ? (spot - strike) : (strike - spot);
  ORC_OPS(1);
  switch (I.in.payoff) {
    case MYQL_PAYOFF_VANILLA:
      return diff > 0.0 ? diff : 0.0;
    case MYQL_PAYOFF_CASH:
      if (mode == MYQL_GREEK_NONE) return diff > 0.0 ? 1.0 : 0.0;
      ORC_OPS(4);
      return 0.5 * std::erfc(-diff * 0.70710678118654752440 / I.eps);
    case MYQL_PAYOFF_ASSET:
      if (mode == MYQL_GREEK_NONE) return diff > 0.0 ? spot : 0.0;
      ORC_OPS(5);
      return spot * 0.5 * std::erfc(-diff * 0.70710678118654752440 / I.eps);
  }
  return 0.0;
}

// Instrument::calculate_to_buffer: base / up / dn payoffs per strike.
static inline void to_buffer(const Instr &I, int mode, const PathState &s, double S0, double h, double *base,
                             double *up, double *dn) {
  const bool bumps = mode != MYQL_GREEK_NONE;
  const double mu = bumps ? (S0 + h) / S0 : 1.0;
  const double md = bumps ? (S0 - h) / S0 : 1.0;
  const double ST = s.logS;
  const size_t nk = I.K.size();
  const int kind = I.in.kind;
  for (size_t i = 0; i < nk; ++i) {
    const double K = I.K[i];
    if (kind == MYQL_INSTR_EUROPEAN) {
      base[i] = payoff(I, mode, ST, K);
      if (bumps) {
        ORC_OPS(2);
        up[i] = payoff(I, mode, ST * mu, K);
        dn[i] = payoff(I, mode, ST * md, K);
      }
    } else if (kind == MYQL_INSTR_GEO_ASIAN || kind == MYQL_INSTR_ARITH_ASIAN ||
               kind == MYQL_INSTR_LOOKBACK) {
      double ex;
      if (kind == MYQL_INSTR_GEO_ASIAN) ex = s.sumLogS;
      else if (kind == MYQL_INSTR_ARITH_ASIAN) ex = s.sumS;
      else {
        const bool call = I.in.option_type == MYQL_CALL;
        ex = I.in.fixed_strike ? (call ? s.maxLogS : s.minLogS) : (call ? s.minLogS : s.maxLogS);
      }
      auto pay = [&](double sp, double e) {
        return I.in.fixed_strike ? payoff(I, mode, e, K) : payoff(I, mode, sp, e);
      };
      base[i] = pay(ST, ex);
      if (bumps) {
        up[i] = pay(ST * mu, ex * mu);
        dn[i] = pay(ST * md, ex * md);
      }
    } else {  // barrier: payoff is evaluated sharp (Mode None)
      auto pay = [&](double sp, int which) {
        double reb = 0.0;
        if (I.in.rebate_timing != MYQL_REBATE_NONE) {
          reb = I.in.rebate;
          if (I.in.rebate_timing == MYQL_REBATE_AT_HIT && s.hit[which])
            reb = I.in.rebate * std::exp(I.in.rebate_rate * (I.in.maturity - s.hit_t[which]));
        }
        const double pv = payoff(I, MYQL_GREEK_NONE, sp, K);
        if (I.in.barrier_action == MYQL_BARRIER_IN) return s.hit[which] ? pv : reb;
        return s.hit[which] ? reb : pv;
      };
      base[i] = pay(ST, 0);
      if (bumps) {
        up[i] = pay(ST * mu, 1);
        dn[i] = pay(ST * md, 2);
      }
    }
  }
}

static inline TrackerCfg tracker_config(const Instr &I, int mode, double S0, double h) {
  TrackerCfg c;
  c.kind = I.in.kind;
  c.up = I.in.barrier_dir == MYQL_BARRIER_UP;
  c.past_avg_log = I.in.past_avg_log;
  c.past_avg_S = I.in.past_avg_S;
  c.t_elapsed = I.in.t_elapsed;
  if (I.in.hist_min_S > 0.0) c.hist_min_logS = std::log(I.in.hist_min_S);
  if (I.in.hist_max_S > 0.0) c.hist_max_logS = std::log(I.in.hist_max_S);
  if (c.kind == MYQL_INSTR_BARRIER) {
    c.b_base = std::log(I.in.barrier);
    if (mode != MYQL_GREEK_NONE) {
      c.b_up = std::log(I.in.barrier * S0 / (S0 + h));
      c.b_dn = std::log(I.in.barrier * S0 / (S0 - h));
    } else {
      c.b_up = c.b_dn = c.b_base;
    }
  }
  return c;
}

// Expected average variance + jump variance (MonteCarloPricer.hpp:31-47,
// ModelParams.hpp:61-79) and the automatic smoothing bandwidth (:51-74).
static inline double expected_avg_variance(const myql_model &m, double T) {
  auto heston = [&](const myql_heston_params &h) {
    if (std::fabs(h.kappa * T) > 1e-8)
      return h.theta + (h.v0 - h.theta) * (1.0 - std::exp(-h.kappa * T)) / (h.kappa * T);
    return h.v0;
  };
  double var = 0.0;
  if (m.n_factors == 0) var += m.vol * m.vol;
  else if (m.n_factors == 1) var += heston(m.heston[0]);
  else var += heston(m.heston[0]) + heston(m.heston[1]);
  if (m.jump_kind == MYQL_JUMP_MERTON)
    var += m.jump[0] * (m.jump[1] * m.jump[1] + m.jump[2] * m.jump[2]);
  else if (m.jump_kind == MYQL_JUMP_KOU)
    var += m.jump[0] * (m.jump[1] / (m.jump[2] * m.jump[2]) + (1.0 - m.jump[1]) / (m.jump[3] * m.jump[3]));
  return var;
}

static inline double auto_eps(const myql_model &m, const Instr &I, int mode, double S0) {
  if (mode == MYQL_GREEK_NONE) return I.in.payoff_eps;
  if (I.in.payoff == MYQL_PAYOFF_VANILLA) return I.in.payoff_eps;
  if (I.in.payoff_eps >= 0.0) return I.in.payoff_eps;
  const double T = I.in.maturity;
  const double sig = std::sqrt(expected_avg_variance(m, T));
  const double scale = I.in.is_strip ? S0 : I.K[0];
  return sig * scale * std::sqrt(T) / 10.0;
}

}  // namespace orc
