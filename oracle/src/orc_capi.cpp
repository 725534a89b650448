// ORACLE (test infrastructure only): CPU FP64 restatement of
// MonteCarloPricer::calculate (reference MonteCarloPricer.hpp:95-481) and
// MultiLevelMonteCarloPricer::calculate (MultiLevelMonteCarloPricer.hpp:46-341),
// exported with plain C entry points for the pytest suite and bench.py's CPU
// baseline. It is never linked into, loaded by or called from the product
// library.
//
// Stream layout: S xoshiro256++ streams; stream s prices the contiguous block
// of paths an OpenMP static schedule over S threads gives thread s (the
// reference's partition at MonteCarloPricer.hpp:257,302). Stream states are
// either jump-ahead (seeded state jumped s * 2^128) or the reference's own
// per-thread seeding seed + s * 999983 (MonteCarloPricer.hpp:191).
// Full-mode CRN: every bumped variant replays the sample's draws from a copy
// of the stream state taken before the base path (MonteCarloPricer.hpp:304).
//
// How each part is pinned (tests/test_oracle.py unless noted):
//   xoshiro256++ next          golden outputs from an independent big-integer
//                              restatement (test_xoshiro_golden)
//   jump-ahead polynomials     equal the published JUMP / LONG_JUMP constants;
//                              brute-force stepping (test_jump_*)
//   splitmix64 seeding         golden vectors of the published splitmix64.
//                              Equality with Boost's xoshiro256pp(u64) seeding:
//                              parity unpinned (Boost is not in the image)
//   uniform                    closed form; forced redraw of 1.0
//   normcdfinv (Acklam)        scipy ndtri, max rel. error 1.2e-9
//   poissinvcdf (Giles)        F(N-1) < U <= F(N) brackets from scipy
//   poisson_inversion          scipy poisson.ppf away from CDF boundaries;
//                              bit equality with Boost: parity unpinned
//   exponential                inversion replaces Boost's ziggurat: parity
//                              unpinned (statistical only: Kou prices vs
//                              Fourier)
//   chi2 quantile, P, Q        scipy golden values to 1e-13 relative
//   NV / NCI / Exact CIR       conditional moments M1-M3 (test_cir_moments)
//   stepper + pricer           Fourier (Gil-Pelaez) prices and Greeks within
//                              4 SE (fixed seeds); barrier in + out = European; Asian
//                              one-step identity and AM-GM; lookback closed
//                              form; in-flight continuity; MLMC vs Fourier
#include <omp.h>

#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "orc_model.hpp"

using namespace orc;

namespace {

thread_local std::string g_err;

const std::vector<std::array<uint64_t, 4>> &jump_table() {
  static std::vector<std::array<uint64_t, 4>> t = jump_polys(64);
  return t;
}

Xoshiro stream_state(uint64_t seed, int layout, uint64_t s) {
  if (layout == MYQL_STREAMS_SEEDED) return seed_state(seed + s * 999983ull);
  Xoshiro g = seed_state(seed);
  const auto &J = jump_table();
  for (int k = 0; k < 64 && (s >> k); ++k)
    if ((s >> k) & 1ull) apply_poly(g, J[k].data());
  return g;
}

// contiguous static partition of n items over S workers
inline void partition(uint64_t n, uint64_t S, uint64_t s, uint64_t &begin, uint64_t &count) {
  const uint64_t q = n / S, rem = n % S;
  count = q + (s < rem ? 1 : 0);
  begin = s * q + (s < rem ? s : rem);
}

struct TableCache {
  std::map<std::pair<uint64_t, size_t>, std::shared_ptr<NciTable>> m;
  bool reference_rows = false;
  // reference row count of the base workspace per d (where the slopes switch
  // from Fritsch-Butland to exact); defaults to the probe's own
  std::map<uint64_t, long> n_fb;
  static uint64_t bits(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
  }
  std::shared_ptr<NciTable> get(const myql_heston_params &p, double dt) {
    // the table depends on (d, n_max) only
    const double d = 4.0 * p.kappa * p.theta / (p.sigma * p.sigma);
    const size_t n_max = nci_rows_nmax(p, dt, reference_rows);
    const uint64_t dbits = bits(d);
    auto fb = n_fb.find(dbits);
    const long nfb = fb != n_fb.end() ? fb->second : static_cast<long>(nci_rows_nmax(p, dt, true));
    auto key = std::make_pair(dbits, n_max);
    auto it = m.find(key);
    if (it != m.end()) return it->second;
    auto tab = build_nci_table(p, dt, reference_rows, nfb);
    m[key] = tab;
    return tab;
  }
};

struct Tables {
  std::shared_ptr<NciTable> t[2];
};

Tables tables_for(TableCache &tc, const myql_model &m, const int32_t schemes[2], double dt) {
  Tables T;
  for (int i = 0; i < m.n_factors; ++i)
    if (schemes[i] == MYQL_SCHEME_NCI) T.t[i] = tc.get(m.heston[i], dt);
  return T;
}

void validate(const myql_model *m, const int32_t *schemes, const myql_instrument *in) {
  if (!m || !schemes || !in) throw Error(MYQL_E_INVALID_ARG, "null argument");
  if (m->n_factors < 0 || m->n_factors > 2) throw Error(MYQL_E_INVALID_ARG, "n_factors must be 0, 1 or 2");
  for (int i = 0; i < m->n_factors; ++i)
    if (schemes[i] < MYQL_SCHEME_NV || schemes[i] > MYQL_SCHEME_EXACT)
      throw Error(MYQL_E_INVALID_ARG, "scheme must be NV, NCI or Exact for each variance factor");
  if (in->n_strikes < 1 || !in->strikes) throw Error(MYQL_E_INVALID_ARG, "instrument needs strikes");
  if (!(in->maturity > 0.0)) throw Error(MYQL_E_INVALID_ARG, "maturity must be positive");
}

enum { ST_PRICE, ST_DELTA, ST_GAMMA, ST_VEGA0, ST_VEGA1, ST_THETA, ST_RHO, ST_VANNA0, ST_VANNA1, ST_CHARM, N_ST };

}  // namespace

extern "C" {

const char *oracle_last_error() { return g_err.c_str(); }

// number of NCI steps that used the exact chi2 quantile instead of the table
uint64_t oracle_nci_fallbacks(int reset) {
  const uint64_t n = nci_fallbacks().load();
  if (reset) nci_fallbacks().store(0);
  return n;
}

// floating-point operations counted since the last reset (builds with
// ORC_COUNT_OPS only; UINT64_MAX otherwise)
uint64_t oracle_op_count(int reset) {
#ifdef ORC_COUNT_OPS
  const uint64_t n = g_ops;
  if (reset) g_ops = 0;
  return n;
#else
  (void)reset;
  return UINT64_MAX;
#endif
}

void oracle_seed_state(uint64_t seed, uint64_t *out) {
  Xoshiro g = seed_state(seed);
  std::memcpy(out, g.s, 32);
}

void oracle_next(uint64_t *state, int64_t n, uint64_t *out) {
  Xoshiro g;
  std::memcpy(g.s, state, 32);
  for (int64_t i = 0; i < n; ++i) out[i] = xnext(g);
  std::memcpy(state, g.s, 32);
}

void oracle_char_poly(uint64_t *out8) {
  Poly P = char_poly();
  std::memcpy(out8, P.w, 64);
}

void oracle_jump_poly(int k, uint64_t *out4) { std::memcpy(out4, jump_table()[k].data(), 32); }

void oracle_apply_poly(uint64_t *state, const uint64_t *poly4) {
  Xoshiro g;
  std::memcpy(g.s, state, 32);
  apply_poly(g, poly4);
  std::memcpy(state, g.s, 32);
}

// x^n mod P for a small exponent n (brute-force cross-check of jump_poly)
void oracle_x_pow_mod(uint64_t n, uint64_t *out4) {
  const Poly P = char_poly();
  Poly r;
  std::memset(&r, 0, sizeof(r));
  r.w[0] = 1;
  for (uint64_t i = 0; i < n; ++i) r = poly_mulx_mod(r, P);
  std::memcpy(out4, r.w, 32);
}

// x^(2^e) mod P (the xoshiro jump / long-jump polynomials for e = 128 / 192)
void oracle_x_pow2_mod(int e, uint64_t *out4) {
  const Poly r = x_pow2_mod(e, char_poly());
  std::memcpy(out4, r.w, 32);
}

void oracle_stream_state(uint64_t seed, int layout, uint64_t s, uint64_t *out4) {
  Xoshiro g = stream_state(seed, layout, s);
  std::memcpy(out4, g.s, 32);
}

double oracle_normcdfinv(double p) { return normcdfinv(p); }
double oracle_poissinvcdf(double lam, double u) { return poissinvcdf(lam, u); }
double oracle_chi2_quantile(double dof, double p) { return chi2_quantile(dof, p); }
double oracle_gamma_p(double a, double x) { return gamma_p(a, x); }
double oracle_gamma_q(double a, double x) { return gamma_q(a, x); }

void oracle_uniforms(uint64_t *state, int64_t n, double *out) {
  Xoshiro g;
  std::memcpy(g.s, state, 32);
  for (int64_t i = 0; i < n; ++i) out[i] = uniform(g);
  std::memcpy(state, g.s, 32);
}

// n jump counts of mean `mean` (jump_count: sequential inversion below 10,
// the exact inverse CDF from 10) from stream seed_state(seed)
void oracle_poisson_samples(double mean, int64_t n, uint64_t seed, int32_t *out) {
  Xoshiro g = seed_state(seed);
  const double em = std::exp(-mean);
  for (int64_t i = 0; i < n; ++i) out[i] = jump_count(mean, em, g);
}

// n independent CIR variance paths of n_steps steps of size dt with one
// scheme, each on its own seeded stream; returns the terminal variances.
int oracle_cir_terminal(int32_t scheme, const myql_heston_params *p, double dt, int64_t n_steps, int64_t n,
                        uint64_t seed, double *out, int nthreads) {
  try {
    std::shared_ptr<NciTable> tab;
    if (scheme == MYQL_SCHEME_NCI) tab = build_nci_table(*p, dt, false);
    Factor f;
    prepare_factor(f, scheme, *p, 0.0, dt, tab);
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      Xoshiro g = seed_state(seed + static_cast<uint64_t>(i) * 999983ull);
      double v = p->v0;
      for (int64_t k = 0; k < n_steps; ++k) v = evolve(f, v, g);
      out[i] = v;
    }
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  }
}

// Trackers and Instrument::calculate_to_buffer on a given log-price path
// logS[0..n_steps] (ASVJStepper::start_path / multi_step / finalize order,
// total time n_steps * dt): base / up / dn payoffs per strike, spot bump
// h = S0 * fd_bump, the instrument's own payoff_eps as the smoothing width.
// Tracker and payoff check only.
int oracle_path_payoffs(const myql_instrument *instr, int32_t greek_mode, double S0, double fd_bump,
                        const double *logS, int64_t n_steps, double dt, double *base, double *up, double *dn) {
  try {
    if (!instr || !logS || !base || n_steps < 1) throw Error(MYQL_E_INVALID_ARG, "bad arguments");
    Instr I;
    I.in = *instr;
    I.K.assign(instr->strikes, instr->strikes + instr->n_strikes);
    I.eps = instr->payoff_eps;
    const double h = S0 * fd_bump;
    const TrackerCfg c = tracker_config(I, greek_mode, S0, h);
    PathState s;
    s.logS = logS[0];
    trk_init(s, c, s.logS, dt);
    const bool pd = path_dependent(c.kind);
    for (int64_t k = 1; k <= n_steps; ++k) {
      s.logS = logS[k];
      if (pd) trk_update(s, c, dt);
    }
    trk_finalize(s, c, dt, static_cast<double>(n_steps) * dt);
    std::vector<double> u(I.K.size()), d(I.K.size());
    to_buffer(I, greek_mode, s, S0, h, base, up ? up : u.data(), dn ? dn : d.data());
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  }
}

// The NCI scheme's table value for row N (the chi2_{d+2N} quantile before
// the scale factor) at the uniforms u[0..n), from the table a pricing call
// with this factor and step builds (reference_rows: the reference's rows and
// exact-inversion rules, else the extended rows), NCI table check only.
int oracle_nci_lookup(const myql_heston_params *p, double dt, int reference_rows, int32_t N, const double *u,
                      int64_t n, double *out) {
  try {
    TableCache tc;
    tc.reference_rows = reference_rows != 0;
    Factor f;
    prepare_factor(f, MYQL_SCHEME_NCI, *p, 0.0, dt, tc.get(*p, dt));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = nci_lookup(f, N, u[i]) / f.K[2];
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  }
}

int oracle_mc_calculate(const myql_model *model, const int32_t *schemes, const myql_instrument *instr,
                        int32_t mode, const myql_mc_config *cfg, double S0, double r, double q,
                        myql_mc_result *res, int nthreads) {
  try {
    validate(model, schemes, instr);
    if (!cfg || !res) throw Error(MYQL_E_INVALID_ARG, "null argument");
    if (cfg->num_paths < 1 || cfg->time_steps < 1) throw Error(MYQL_E_INVALID_ARG, "num_paths and time_steps must be >= 1");
    // quasi-Monte Carlo points are checked through oracle/sobol.py, not here
    if (cfg->stream_layout == MYQL_STREAMS_SOBOL) throw Error(MYQL_E_UNSUPPORTED, "the oracle draws pseudo-random streams only");
    const myql_model &M0 = *model;
    Instr I;
    I.in = *instr;
    I.K.assign(instr->strikes, instr->strikes + instr->n_strikes);
    I.eps = auto_eps(M0, I, mode, S0);
    const double T = I.in.maturity;
    const uint64_t steps = cfg->time_steps, M = cfg->num_paths;
    const double dt = T / static_cast<double>(steps);
    const double h = S0 * cfg->fd_bump;
    const size_t nk = I.K.size();
    const bool pd = path_dependent(I.in.kind);
    const int nf = M0.n_factors;
    const bool full = mode == MYQL_GREEK_FULL, ess = mode != MYQL_GREEK_NONE;
    const TrackerCfg tc = tracker_config(I, mode, S0, h);

    TableCache cache;
    cache.reference_rows = (cfg->flags & MYQL_FLAG_REFERENCE_NCI_ROWS) != 0;
    for (int i = 0; i < M0.n_factors; ++i)
      if (schemes[i] == MYQL_SCHEME_NCI) {
        const myql_heston_params &h = M0.heston[i];
        cache.n_fb[TableCache::bits(4.0 * h.kappa * h.theta / (h.sigma * h.sigma))] =
            static_cast<long>(nci_rows_nmax(h, dt, true));
      }
    auto mk = [&](const myql_model &m, double dtv, double rv, double Tv) {
      Tables tb = tables_for(cache, m, schemes, dtv);
      return make_stepper(m, schemes, dtv, rv, q, Tv, pd, tb.t[0], tb.t[1]);
    };
    auto bump = [&](int idx, double dv) {
      myql_model m = M0;
      if (nf == 0) m.vol += dv;
      else m.heston[idx].v0 += dv;
      return m;
    };
    const Stepper base = mk(M0, dt, r, T);
    Stepper vup[2], vdn[2], Tup, Tdn, rup, rdn;
    double cv[2] = {1.0, 1.0};
    const double Tu = T + cfg->t_bump, Td = T - cfg->t_bump;
    if (full) {
      for (int i = 0; i < std::max(nf, 1) && i < 2; ++i) {
        vup[i] = mk(bump(i, cfg->vol_bump), dt, r, T);
        vdn[i] = mk(bump(i, -cfg->vol_bump), dt, r, T);
        if (nf > 0) cv[i] = 2.0 * std::sqrt(M0.heston[i].v0);
      }
      Tup = mk(M0, Tu / steps, r, Tu);
      Tdn = mk(M0, Td / steps, r, Td);
      rup = mk(M0, dt, r + cfg->r_bump, T);
      rdn = mk(M0, dt, r - cfg->r_bump, T);
    }
    const int nvega = std::max(nf, 1);  // 0-factor bumps vol once
    const double inv2dv = 1.0 / (2.0 * cfg->vol_bump), inv2dT = 1.0 / (2.0 * cfg->t_bump),
                 inv2dr = 1.0 / (2.0 * cfg->r_bump);
    const double dfTu = std::exp(-r * cfg->t_bump), dfTd = std::exp(r * cfg->t_bump);
    const double dfru = std::exp(-cfg->r_bump * T), dfrd = std::exp(cfg->r_bump * T);

    uint64_t S = cfg->num_streams ? cfg->num_streams : std::min<uint64_t>(65536, M);
    if (S > M) S = M;
    const int layout = cfg->stream_layout;
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<std::vector<double>> partial(nt, std::vector<double>(2 * N_ST * nk, 0.0));

#pragma omp parallel num_threads(nt)
    {
      const int tid = omp_get_thread_num();
      std::vector<double> &acc = partial[tid];
      std::vector<double> pb(nk), pu(nk), pdn(nk), d1(nk), d2(nk);
      std::vector<double> a_p(nk), a_u(nk), a_d(nk), b_p(nk), b_u(nk), b_d(nk);
      auto add = [&](int st, size_t i, double x) {
        acc[(2 * st) * nk + i] += x;
        acc[(2 * st + 1) * nk + i] += x * x;
        ORC_OPS(3);
      };
      auto run = [&](const Stepper &stp, Xoshiro &g, double *p, double *u, double *d) {
        PathState ps;
        start_path(stp, ps, tc, S0);
        multi_step(stp, ps, tc, g, static_cast<unsigned>(steps));
        to_buffer(I, mode, ps, S0, h, p, u, d);
      };
#pragma omp for schedule(static)
      for (int64_t s = 0; s < static_cast<int64_t>(S); ++s) {
        uint64_t b0, cnt;
        partition(M, S, s, b0, cnt);
        Xoshiro g = stream_state(cfg->seed, layout, s);
        for (uint64_t pth = 0; pth < cnt; ++pth) {
          const Xoshiro snap = g;
          run(base, g, pb.data(), pu.data(), pdn.data());
          for (size_t i = 0; i < nk; ++i) {
            add(ST_PRICE, i, pb[i]);
            if (ess) {
              add(ST_DELTA, i, (pu[i] - pdn[i]) * (1.0 / (2.0 * h)));
              add(ST_GAMMA, i, (pu[i] - pb[i] * 2.0 + pdn[i]) * (1.0 / (h * h)));
              ORC_OPS(6);
            }
          }
          if (!full) continue;
          for (int f = 0; f < nvega; ++f) {
            Xoshiro g1 = snap, g2 = snap;
            run(vup[f], g1, a_p.data(), a_u.data(), a_d.data());
            run(vdn[f], g2, b_p.data(), b_u.data(), b_d.data());
            for (size_t i = 0; i < nk; ++i) {
              add(ST_VEGA0 + f, i, (a_p[i] - b_p[i]) * inv2dv * cv[f]);
              add(ST_VANNA0 + f, i, ((a_u[i] - a_d[i]) - (b_u[i] - b_d[i])) * (inv2dv / (2.0 * h)) * cv[f]);
              ORC_OPS(8);
            }
          }
          {
            Xoshiro g1 = snap, g2 = snap;
            run(Tup, g1, a_p.data(), a_u.data(), a_d.data());
            run(Tdn, g2, b_p.data(), b_u.data(), b_d.data());
            for (size_t i = 0; i < nk; ++i) {
              add(ST_THETA, i, (a_p[i] * dfTu - b_p[i] * dfTd) * inv2dT);
              add(ST_CHARM, i, ((a_u[i] - a_d[i]) * dfTu - (b_u[i] - b_d[i]) * dfTd) * (inv2dT / (2.0 * h)));
              ORC_OPS(11);
            }
          }
          {
            Xoshiro g1 = snap, g2 = snap;
            run(rup, g1, a_p.data(), a_u.data(), a_d.data());
            run(rdn, g2, b_p.data(), b_u.data(), b_d.data());
            for (size_t i = 0; i < nk; ++i) {
              add(ST_RHO, i, (a_p[i] * dfru - b_p[i] * dfrd) * inv2dr);
              ORC_OPS(4);
            }
          }
        }
      }
    }
    std::vector<double> tot(2 * N_ST * nk, 0.0);
    for (int t = 0; t < nt; ++t)
      for (size_t j = 0; j < tot.size(); ++j) tot[j] += partial[t][j];

    const double df = std::exp(-r * T), invM = 1.0 / static_cast<double>(M);
    auto stats = [&](int st, double *val, double *err) {
      for (size_t i = 0; i < nk; ++i) {
        const double mean = tot[(2 * st) * nk + i] * invM;
        const double var = tot[(2 * st + 1) * nk + i] * invM - mean * mean;
        if (val) val[i] = mean * df;
        if (err) err[i] = std::sqrt(var * invM) * df;
      }
    };
    stats(ST_PRICE, res->price, res->price_std_err);
    if (ess) {
      stats(ST_DELTA, res->delta, res->delta_std_err);
      stats(ST_GAMMA, res->gamma, res->gamma_std_err);
    }
    if (full) {
      stats(ST_VEGA0, res->vega[0], res->vega_std_err[0]);
      stats(ST_VEGA1, res->vega[1], res->vega_std_err[1]);
      stats(ST_THETA, res->theta, res->theta_std_err);
      stats(ST_RHO, res->rho, res->rho_std_err);
      stats(ST_VANNA0, res->vanna[0], res->vanna_std_err[0]);
      stats(ST_VANNA1, res->vanna[1], res->vanna_std_err[1]);
      stats(ST_CHARM, res->charm, res->charm_std_err);
    }
    const uint64_t variants = full ? (nf == 2 ? 9 : 7) : 1;
    res->total_steps = M * steps * variants;
    res->device_ms = 0.0;
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    return MYQL_E_INVALID_ARG;
  }
}

// ---------------------------------------------------------------------------
// Debug exports (tests only): one path of the plain stepper or of the MLMC
// coupled stepper from the state seed_state(seed), with its per-step record
// (multi_step / multi_step_coupled `rec`), the final tracker states and the
// log of every uniform drawn, so that a test can replay the path's draws
// through the step equations.
// ---------------------------------------------------------------------------
namespace {
constexpr int FIN = 10;  // S (finalized), V0, V1, sumLogS, sumS, min, max, hit0..2

void put_final(const PathState &s, double *o) {
  const double v[FIN] = {s.logS, s.V[0], s.V[1], s.sumLogS, s.sumS, s.minLogS, s.maxLogS,
                         s.hit[0] ? 1.0 : 0.0, s.hit[1] ? 1.0 : 0.0, s.hit[2] ? 1.0 : 0.0};
  for (int i = 0; i < FIN; ++i) o[i] = v[i];
}

struct DrawLog {
  std::vector<double> log;
#ifndef ORC_NO_DRAW_LOG
  DrawLog() { g_draw_log = &log; }
  ~DrawLog() { g_draw_log = nullptr; }
#else
  DrawLog() { throw Error(MYQL_E_UNSUPPORTED, "this oracle build has no draw log"); }
#endif
  void copy(double *draws, int64_t max_draws, int64_t *n_draws) const {
    *n_draws = static_cast<int64_t>(log.size());
    const size_t n = std::min<size_t>(log.size(), static_cast<size_t>(std::max<int64_t>(max_draws, 0)));
    std::copy(log.begin(), log.begin() + n, draws);
  }
};
}  // namespace

int oracle_debug_plain_path(const myql_model *model, const int32_t *schemes, const myql_instrument *instr,
                            uint64_t steps, uint64_t seed, double S0, double r, double q, double *rec,
                            double *fin, double *draws, int64_t max_draws, int64_t *n_draws) {
  try {
    validate(model, schemes, instr);
    if (!rec || !fin || !draws || !n_draws || steps < 1) throw Error(MYQL_E_INVALID_ARG, "bad arguments");
    Instr I;
    I.in = *instr;
    I.K.assign(instr->strikes, instr->strikes + 1);
    const double T = I.in.maturity, dt = T / static_cast<double>(steps);
    const bool pd = path_dependent(I.in.kind);
    const TrackerCfg tc = tracker_config(I, MYQL_GREEK_NONE, S0, S0 * 1e-4);
    TableCache cache;
    Tables tb = tables_for(cache, *model, schemes, dt);
    const Stepper st = make_stepper(*model, schemes, dt, r, q, T, pd, tb.t[0], tb.t[1]);
    Xoshiro g = seed_state(seed);
    DrawLog dl;
    PathState s;
    start_path(st, s, tc, S0);
    multi_step(st, s, tc, g, static_cast<unsigned>(steps), rec);
    put_final(s, fin);
    dl.copy(draws, max_draws, n_draws);
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  }
}

int oracle_debug_coupled_path(const myql_model *model, const int32_t *schemes, const myql_instrument *instr,
                              uint64_t coarse_steps, uint64_t seed, double S0, double r, double q, double *rec,
                              double *fin_f, double *fin_c, double *draws, int64_t max_draws, int64_t *n_draws) {
  try {
    validate(model, schemes, instr);
    if (!rec || !fin_f || !fin_c || !draws || !n_draws || coarse_steps < 1)
      throw Error(MYQL_E_INVALID_ARG, "bad arguments");
    Instr I;
    I.in = *instr;
    I.K.assign(instr->strikes, instr->strikes + 1);
    const double T = I.in.maturity, dtf = T / static_cast<double>(2 * coarse_steps);
    const bool pd = path_dependent(I.in.kind);
    const TrackerCfg tc = tracker_config(I, MYQL_GREEK_NONE, S0, S0 * 1e-4);
    TableCache cache;
    Tables tb = tables_for(cache, *model, schemes, dtf);
    const Stepper F = make_stepper(*model, schemes, dtf, r, q, T, pd, tb.t[0], tb.t[1]);
    const Stepper C = make_stepper(*model, schemes, 2.0 * dtf, r, q, T, pd, tb.t[0], tb.t[1]);
    Xoshiro g = seed_state(seed);
    DrawLog dl;
    PathState sf, sc;
    start_path(F, sf, tc, S0);
    start_path(C, sc, tc, S0);
    multi_step_coupled(F, C, sf, sc, tc, g, static_cast<unsigned>(coarse_steps), rec);
    put_final(sf, fin_f);
    put_final(sc, fin_c);
    dl.copy(draws, max_draws, n_draws);
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  }
}

// Weights of the level means 0..L (the multilevel estimator's extension
// beyond the reference, which sums the means with weight 1,
// MultiLevelMonteCarloPricer.hpp:215-235). ML2R (Lemaire & Pages 2017):
// the Richardson weights w solve the Vandermonde system
// sum_j w_j x_j^k = [k == 0], k = 0..L, x_j = 2^(-alpha j), here by Gaussian
// elimination with partial pivoting; level l carries W_l = sum_{j >= l} w_j.
static std::vector<double> level_weights(size_t L, double alpha, bool ml2r) {
  std::vector<double> W(L + 1, 1.0);
  if (!ml2r) return W;
  const size_t n = L + 1;
  std::vector<double> A(n * n), b(n, 0.0);
  for (size_t k = 0; k < n; ++k)
    for (size_t j = 0; j < n; ++j) A[k * n + j] = std::pow(2.0, -alpha * static_cast<double>(j * k));
  b[0] = 1.0;
  for (size_t c = 0; c < n; ++c) {
    size_t piv = c;
    for (size_t r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    for (size_t j = 0; j < n; ++j) std::swap(A[c * n + j], A[piv * n + j]);
    std::swap(b[c], b[piv]);
    for (size_t r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (size_t j = c; j < n; ++j) A[r * n + j] -= f * A[c * n + j];
      b[r] -= f * b[c];
    }
  }
  std::vector<double> w(n);
  for (size_t c = n; c-- > 0;) {
    double t = b[c];
    for (size_t j = c + 1; j < n; ++j) t -= A[c * n + j] * w[j];
    w[c] = t / A[c * n + c];
  }
  double t = 0.0;
  for (size_t l = n; l-- > 0;) {
    t += w[l];
    W[l] = t;
  }
  W[0] = 1.0;
  return W;
}

int oracle_mlmc_level_weights(int64_t L, double alpha, int32_t ml2r, double *W) {
  if (L < 0 || L > 30 || !W) return MYQL_E_INVALID_ARG;
  const std::vector<double> w = level_weights(static_cast<size_t>(L), alpha, ml2r != 0);
  for (int64_t l = 0; l <= L; ++l) W[l] = w[l];
  return MYQL_OK;
}

int oracle_mlmc_calculate(const myql_model *model, const int32_t *schemes, const myql_instrument *instr,
                          int32_t mode, const myql_mlmc_config *cfg, double S0, double r, double q,
                          myql_mlmc_result *res, int nthreads) {
  try {
    validate(model, schemes, instr);
    if (!cfg || !res) throw Error(MYQL_E_INVALID_ARG, "null argument");
    if (cfg->estimator != MYQL_MLMC_GILES && cfg->estimator != MYQL_MLMC_ML2R)
      throw Error(MYQL_E_INVALID_ARG, "bad estimator");
    const myql_model &M0 = *model;
    Instr I;
    I.in = *instr;
    I.K.assign(instr->strikes, instr->strikes + 1);
    I.eps = I.in.payoff_eps;
    const double T = I.in.maturity;
    const bool pd = path_dependent(I.in.kind);
    const double h = S0 * cfg->fd_bump;
    const TrackerCfg tc = tracker_config(I, mode, S0, h);
    const size_t Lmax = cfg->L_max;
    const bool greeks = mode != MYQL_GREEK_NONE;
    const bool ml2r = cfg->estimator == MYQL_MLMC_ML2R;
    uint64_t S = cfg->num_streams ? cfg->num_streams : 65536;
    std::vector<Xoshiro> pool(S);
    for (uint64_t s = 0; s < S; ++s) pool[s] = stream_state(cfg->seed, cfg->stream_layout, s);
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    TableCache cache;
    cache.reference_rows = (cfg->flags & MYQL_FLAG_REFERENCE_NCI_ROWS) != 0;
    constexpr int NS = 6;  // price, delta and gamma differences, each with its square

    // per path: the level's price difference and, with Greeks, the
    // differences of the fine and coarse central differences of the
    // spot-bumped payoffs (calculate_to_buffer's up / dn buffers)
    auto add = [&](double *a, double p, double u, double d) {
      a[0] += p;
      a[1] += p * p;
      if (greeks) {
        a[2] += u;
        a[3] += u * u;
        a[4] += d;
        a[5] += d * d;
      }
    };
    auto evaluate_level = [&](size_t l, uint64_t dN, double *acc) {
      std::vector<double> part(S * NS, 0.0);
      if (l == 0) {
        const uint64_t cs = cfg->base_steps;
        const double dtc = T / static_cast<double>(cs);
        Tables tb = tables_for(cache, M0, schemes, dtc);
        const Stepper st = make_stepper(M0, schemes, dtc, r, q, T, pd, tb.t[0], tb.t[1]);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t s = 0; s < static_cast<int64_t>(S); ++s) {
          uint64_t b0, cnt;
          partition(dN, S, s, b0, cnt);
          double p0, d1, d2;
          for (uint64_t i = 0; i < cnt; ++i) {
            PathState st0;
            start_path(st, st0, tc, S0);
            multi_step(st, st0, tc, pool[s], static_cast<unsigned>(cs));
            to_buffer(I, mode, st0, S0, h, &p0, &d1, &d2);
            add(&part[s * NS], p0, (d1 - d2) / (2.0 * h), (d1 - 2.0 * p0 + d2) / (h * h));
          }
        }
      } else {
        const uint64_t fs = cfg->base_steps * (1ull << l), cs = fs / 2;
        const double dtf = T / static_cast<double>(fs);
        Tables tb = tables_for(cache, M0, schemes, dtf);
        const Stepper F = make_stepper(M0, schemes, dtf, r, q, T, pd, tb.t[0], tb.t[1]);
        const Stepper C = make_stepper(M0, schemes, 2.0 * dtf, r, q, T, pd, tb.t[0], tb.t[1]);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t s = 0; s < static_cast<int64_t>(S); ++s) {
          uint64_t b0, cnt;
          partition(dN, S, s, b0, cnt);
          double pf, pc, fu, fd, cu, cd;
          for (uint64_t i = 0; i < cnt; ++i) {
            PathState sf, sc;
            start_path(F, sf, tc, S0);
            start_path(C, sc, tc, S0);
            multi_step_coupled(F, C, sf, sc, tc, pool[s], static_cast<unsigned>(cs));
            to_buffer(I, mode, sf, S0, h, &pf, &fu, &fd);
            to_buffer(I, mode, sc, S0, h, &pc, &cu, &cd);
            add(&part[s * NS], pf - pc, (fu - fd) / (2.0 * h) - (cu - cd) / (2.0 * h),
                (fu - 2.0 * pf + fd) / (h * h) - (cu - 2.0 * pc + cd) / (h * h));
          }
        }
      }
      for (uint64_t s = 0; s < S; ++s)
        for (int a = 0; a < NS; ++a) acc[a] += part[s * NS + a];
    };

    size_t L = cfg->L_min;
    std::vector<uint64_t> N(Lmax + 1, 0);
    std::vector<std::array<double, NS>> acc(Lmax + 1);
    for (auto &a : acc) a.fill(0.0);
    std::vector<double> V(Lmax + 1, 0.0), cost(Lmax + 1, 0.0);
    for (size_t l = 0; l <= Lmax; ++l) cost[l] = static_cast<double>(cfg->base_steps * (1ull << l));
    auto mean = [&](size_t l, int a) { return acc[l][a] / static_cast<double>(N[l]); };
    auto var = [&](size_t l, int a) {
      const double m = mean(l, a);
      return std::max(1e-18, acc[l][a + 1] / static_cast<double>(N[l]) - m * m);
    };
    double alpha = cfg->alpha, beta = cfg->beta;
    const double eps = cfg->epsilon;
    std::vector<double> W = level_weights(L, cfg->alpha, ml2r);
    // least-cost sample sizes of sum_l W_l mean_l at variance eps^2 / 2
    auto optimal = [&](size_t l, size_t Lc) -> uint64_t {
      if (!(V[l] > 1e-18)) return 0;
      double sumVC = 0.0;
      for (size_t k = 0; k <= Lc; ++k) sumVC += std::fabs(W[k]) * std::sqrt(V[k] * cost[k]);
      return static_cast<uint64_t>(
          std::ceil(2.0 * std::pow(eps, -2.0) * std::fabs(W[l]) * std::sqrt(V[l] / cost[l]) * sumVC));
    };
    bool converged = false;
    while (!converged) {
      W = level_weights(L, cfg->alpha, ml2r);
      for (size_t l = 0; l <= L; ++l) {
        const uint64_t opt = optimal(l, L);
        const uint64_t pilot = std::max<uint64_t>(20, cfg->N0 / (1ull << l));
        const uint64_t target = std::max(pilot, opt);
        if (N[l] < target) {
          const uint64_t dN = target - N[l];
          evaluate_level(l, dN, acc[l].data());
          N[l] += dN;
          if (N[l] > 1) V[l] = var(l, 0);
        }
      }
      if (cfg->adaptive_rates && L >= 2) {
        auto order = [&](const std::vector<double> &vals) {
          double sl = 0, sy = 0, sly = 0, sl2 = 0;
          int cnt = 0;
          const size_t l0 = L - 2;
          for (size_t l = l0; l <= L; ++l) {
            if (vals[l] > 1e-16) {
              const double y = std::log(vals[l]);
              const double x = static_cast<double>(l - l0);
              sl += x;
              sy += y;
              sly += x * y;
              sl2 += x * x;
              ++cnt;
            }
          }
          if (cnt < 2) return -1.0;
          const double slope = (cnt * sly - sl * sy) / (cnt * sl2 - sl * sl);
          return -slope / std::log(2.0);
        };
        std::vector<double> am(L + 1);
        for (size_t l = 0; l <= L; ++l) am[l] = std::fabs(mean(l, 0));
        const double ea = order(am), eb = order(V);
        if (ea > 0.3) alpha = std::min(2.0, ea);
        if (eb > 0.3) beta = std::min(2.0, eb);
      }
      bool more = false;
      for (size_t l = 0; l <= L; ++l)
        if (optimal(l, L) > N[l]) more = true;
      if (more) continue;
      double bias;
      if (!ml2r) {
        double eff = std::fabs(mean(L, 0));
        if (L > 0) eff = std::max(eff, 0.5 * std::fabs(mean(L - 1, 0)));
        bias = eff / (std::pow(2.0, alpha) - 1.0);
      } else if (L == 0) {
        bias = HUGE_VAL;
      } else {
        // |estimate at depth L - estimate at depth L - 1|
        const std::vector<double> Wp = level_weights(L - 1, cfg->alpha, true);
        double a = 0.0, b = 0.0;
        for (size_t l = 0; l <= L; ++l) a += W[l] * mean(l, 0);
        for (size_t l = 0; l < L; ++l) b += Wp[l] * mean(l, 0);
        bias = std::fabs(a - b);
      }
      if (bias > eps / 1.4142135623730951 && L < Lmax) ++L;
      else converged = true;
    }
    W = level_weights(L, cfg->alpha, ml2r);
    const double df = std::exp(-r * T);
    double est[3] = {0.0, 0.0, 0.0}, ev[3] = {0.0, 0.0, 0.0};
    res->total_steps = 0;
    res->alpha_estimated = alpha;
    res->beta_estimated = beta;
    res->n_levels = static_cast<int32_t>(L + 1);
    for (size_t l = 0; l <= L; ++l) {
      const uint64_t spp = (l == 0) ? cfg->base_steps : (cfg->base_steps * (1ull << l) * 3 / 2);
      res->total_steps += N[l] * spp;
      if (res->paths_per_level) res->paths_per_level[l] = N[l];
      if (res->mean_diffs) res->mean_diffs[l] = mean(l, 0) * df;
      if (res->var_diffs) res->var_diffs[l] = V[l] * df * df;
      if (res->level_weights) res->level_weights[l] = W[l];
      for (int a = 0; a < (greeks ? 3 : 1); ++a) {
        est[a] += W[l] * mean(l, 2 * a);
        ev[a] += W[l] * W[l] * (a == 0 ? V[l] : var(l, 2 * a)) / static_cast<double>(N[l]);
      }
    }
    res->price = est[0] * df;
    res->price_std_err = std::sqrt(ev[0]) * df;
    res->delta = est[1] * df;
    res->delta_std_err = std::sqrt(ev[1]) * df;
    res->gamma = est[2] * df;
    res->gamma_std_err = std::sqrt(ev[2]) * df;
    res->device_ms = 0.0;
    return MYQL_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    return MYQL_E_INVALID_ARG;
  }
}

}  // extern "C"
