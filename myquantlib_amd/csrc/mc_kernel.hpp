// Path-simulation kernels of the MI355X Monte Carlo engine.
//
// One lane owns one xoshiro256++ stream and walks the samples of that stream
// in order (the OpenMP static partition the reference uses,
// MonteCarloPricer.hpp:257,302). A sample's Greek variants either run in
// lockstep in the lane's registers, sharing every draw (NV/NCI schemes:
// fixed draw counts), or are replayed one after another from a register
// copy of the stream state (Exact scheme, path-dependent jumps) -- both give
// exactly the reference's common-random-number replay (:304, :328-405).
//
// Per block of 256 lanes (4 wavefronts) the terminal state of every variant
// is staged in LDS; the strike loop then maps lanes to strikes, so each lane
// keeps the FP64 sum / sum-of-squares of one strike in registers and no
// atomics are used. Work units (job, 256 streams) are pulled from a queue by
// a persistent grid; each unit writes its own partial sums, which a second
// kernel adds in unit order, so results do not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_math.hpp"
#include "engine_types.hpp"

namespace myql {

template <typename R>
__device__ __forceinline__ R dsqrt(R x);
template <>
__device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }  // v_sqrt_f32, 1 ulp

template <typename R>
__device__ __forceinline__ R dexp(R x);
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }

template <typename R>
__device__ __forceinline__ R dlog(R x);
template <>
__device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <>
__device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }

// sqrt in FP64 from the FP32 reciprocal square root and one Newton step
// (relative error ~5e-15; the library sqrt's scaling and rounding steps cost
// three times as many instructions)
__device__ __forceinline__ double sqrt_nr(double a) {
  if (!(a > 1e-30)) return a > 0.0 ? sqrt(a) : 0.0;
  const double y = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(a)));
  return a * (y * (1.5 - 0.5 * a * y * y));
}

// Lockstep groups of Full-mode Europeans in FP32 kernels keep the two
// maturity-bumped variants (T+ in slot 1 + 2 max(NF, 1), T- next; the
// reference's T +- t_bump replays, MonteCarloPricer.hpp:374-395) in FP64:
// theta and charm difference those two paths and divide by 2 t_bump
// (2e-4), which FP32 rounding of the log-price would not survive. Their NCI
// variance steps read the FP64 chi-squared rows (Job::nci64), so they take
// the FP64 kernel's Poisson counts too. Their
// terminal spots keep FP64 precision through the FP32 stage: T+ as a
// hi + lo pair of floats (lo in the r+ variant's slot, unused by Europeans,
// whose rate-bumped paths are derived from the base path), T- relative to
// it as e^(x_T- - x_T+) - 1.
template <typename R, int NV, int NFX, bool PD>
constexpr bool t_slots_fp64() { return sizeof(R) == 4 && !PD && NV == 3 + 2 * NFX; }
template <int NFX>
constexpr int t_slot0() { return 1 + 2 * NFX; }
// whether a kernel's paths carry the log-return instead of the log-price
template <typename R, bool PD>
constexpr bool log_return_paths() { return sizeof(R) == 4 && !PD; }

template <typename R>
struct Trk {
  R sumlog, sumS, mn, mx, t;
  int hits;  // bit 0 base barrier, bit 1 barrier for S0+h, bit 2 for S0-h
  R ht[3];
};

// ---- trackers (PathTrackers.hpp:8-300) -----------------------------------
template <typename R>
__device__ __forceinline__ void trk_init(const Job<R> *__restrict__ J, Trk<R> &s, R x, R dt) {
  s.sumlog = 0;
  s.sumS = 0;
  s.mn = x;
  s.mx = x;
  s.t = 0;
  s.hits = 0;
  s.ht[0] = s.ht[1] = s.ht[2] = 0;
  switch (J->tracker) {
    case 1:  // geometric Asian
      s.sumlog = static_cast<R>(J->init_sumlog) + R(0.5) * x * dt;
      break;
    case 2:  // arithmetic Asian
      s.sumlog = static_cast<R>(J->init_sumlog) + R(0.5) * x * dt;
      s.sumS = static_cast<R>(J->init_sumS) + R(0.5) * dexp<R>(x) * dt;
      break;
    case 3:  // lookback, optionally seeded by the observed history
      if (J->hist_min_logS < 1e300) s.mn = static_cast<R>(J->hist_min_logS);
      if (J->hist_max_logS > -1e300) s.mx = static_cast<R>(J->hist_max_logS);
      break;
    case 4:  // barrier: day-0 breaches
      if (J->barrier_up) {
        s.hits = (x >= J->b_base ? 1 : 0) | (x >= J->b_up ? 2 : 0) | (x >= J->b_dn ? 4 : 0);
      } else {
        s.hits = (x <= J->b_base ? 1 : 0) | (x <= J->b_up ? 2 : 0) | (x <= J->b_dn ? 4 : 0);
      }
      break;
  }
  if (J->fuse_asian) {  // a fused arithmetic Asian's state, as case 2
    s.sumlog = static_cast<R>(J->init_sumlog) + R(0.5) * x * dt;
    s.sumS = static_cast<R>(J->init_sumS) + R(0.5) * dexp<R>(x) * dt;
  }
}

template <typename R>
__device__ __forceinline__ void mark(Trk<R> &s, int bit, int idx) {
  if (!(s.hits & bit)) {
    s.hits |= bit;
    s.ht[idx] = s.t;
  }
}

template <typename R>
__device__ __forceinline__ void trk_update(const Job<R> *__restrict__ J, Trk<R> &s, R x, R dt) {
  switch (J->tracker) {
    case 1:
      s.sumlog += x * dt;
      break;
    case 2:
      s.sumlog += x * dt;
      s.sumS += dexp<R>(x) * dt;
      break;
    case 3:
      s.mx = (x > s.mx) ? x : s.mx;
      s.mn = (x < s.mn) ? x : s.mn;
      break;
    case 4:
      s.t += dt;
      // nested tests, easiest barrier first
      if (J->barrier_up) {
        if (x >= J->b_up) {
          mark(s, 2, 1);
          if (x >= J->b_base) {
            mark(s, 1, 0);
            if (x >= J->b_dn) mark(s, 4, 2);
          }
        }
      } else {
        if (x <= J->b_dn) {
          mark(s, 4, 2);
          if (x <= J->b_base) {
            mark(s, 1, 0);
            if (x <= J->b_up) mark(s, 2, 1);
          }
        }
      }
      break;
  }
  if (J->fuse_asian) {
    s.sumlog += x * dt;
    s.sumS += dexp<R>(x) * dt;
  }
}

// Finalize and stage one variant's terminal data in LDS.
template <typename R, int BLK = BLOCK>
__device__ __forceinline__ void finalize_stage(const Job<R> *__restrict__ J, const Trk<R> &s, R x, R dt,
                                               R T, R *__restrict__ stage, int v, int tid) {
  R aux = 0;
  switch (J->tracker) {
    case 1: {
      const R sl = s.sumlog - R(0.5) * x * dt;
      aux = dexp<R>(sl / (T + static_cast<R>(J->t_elapsed)));
      break;
    }
    case 2: {
      const R sS = s.sumS - R(0.5) * dexp<R>(x) * dt;
      aux = sS / (T + static_cast<R>(J->t_elapsed));
      break;
    }
    case 3: {
      const R e = J->fixed_strike ? (J->call ? s.mx : s.mn) : (J->call ? s.mn : s.mx);
      aux = dexp<R>(e);
      break;
    }
  }
  if (J->fuse_asian) {
    const R sS = s.sumS - R(0.5) * dexp<R>(x) * dt;
    aux = sS / (T + static_cast<R>(J->t_elapsed));
  }
  stage[(F_SPOT * MAXV + v) * BLK + tid] = dexp<R>(x);
  if (J->n_fields > F_AUX) stage[(F_AUX * MAXV + v) * BLK + tid] = aux;
  if (J->n_fields > F_FLAGS) stage[(F_FLAGS * MAXV + v) * BLK + tid] = static_cast<R>(s.hits);
  if (J->n_fields > F_HT0) {
    stage[(F_HT0 * MAXV + v) * BLK + tid] = s.ht[0];
    stage[(F_HT1 * MAXV + v) * BLK + tid] = s.ht[1];
    stage[(F_HT2 * MAXV + v) * BLK + tid] = s.ht[2];
  }
}

// ---- gamma variates of SchemeExact (VolSchemes.hpp:129-203) ---------------
template <typename R, typename Gen>
__device__ static R gamma_draw(double alpha_base, int N, Gen &g) {
  const double alpha = alpha_base + N;
  if (alpha > 1.0) {
    const double dd = alpha - 1.0 / 3.0;
    const R d = static_cast<R>(dd);
    const R c = static_cast<R>(1.0 / sqrt(9.0 * dd));
    for (;;) {
      R Z, v;
      do {
        Z = normcdfinv<R>(uniform(g));
        v = R(1) + c * Z;
      } while (v <= R(0));
      v = v * v * v;
      const R U = static_cast<R>(uniform(g));
      if (U < R(1) - R(0.0331) * Z * Z * Z * Z) return d * v;
      if (dlog<R>(U) < R(0.5) * Z * Z + d * (R(1) - v + dlog<R>(v))) return d * v;
    }
  }
  // Ahrens-Dieter GS for alpha <= 1: one uniform picks the piece and is
  // reused, rescaled, as the acceptance uniform
  const double e = 2.71828182845904523536;
  const R p = static_cast<R>(e / (alpha_base + e));
  const R inv_p = static_cast<R>((alpha_base + e) / e);
  const R one_m_p = R(1) - p;
  const R one_m_a = static_cast<R>(1.0 - alpha_base);
  const R inv_a = static_cast<R>(1.0 / alpha_base);
  for (;;) {
    const R u = static_cast<R>(uniform(g));
    const R y = exponential<R>(R(1), g);
    if (u < p) {
      const R x = dexp<R>(-y * inv_a);
      if (u * inv_p < R(1) - x) return x;
      if (u < p * dexp<R>(-x)) return x;
    } else {
      const R x = R(1) + y;
      const R inv_h = R(1) / (R(1) + one_m_a * y);
      if (u < p + one_m_p * inv_h) return x;
      if (u < p + one_m_p * dexp<R>(-one_m_a * dlog<R>(x))) return x;
    }
  }
}

// ---- one lane-local log jump (JumpPolicies.hpp:55-135) ---------------------
template <typename R, typename Gen>
__device__ __forceinline__ R log_jump(const Job<R> *__restrict__ J, const VarParams<R> *__restrict__ P,
                                      Gen &g) {
  const int n = jump_count(P->jmean, P->jexpm, uniform(g));
  if (n == 0) return P->jdrift;
  if (J->jump_kind == 1) {
    const R Z = normcdfinv<R>(uniform(g));
    return P->jdrift + static_cast<R>(n) * J->jmu + dsqrt<R>(static_cast<R>(n)) * J->jdelta * Z;
  }
  R tot = 0;
  for (int i = 0; i < n; ++i) {
    const double u = uniform(g);
    if (u < static_cast<double>(J->kp)) tot += exponential<R>(J->keta1, g);
    else tot -= exponential<R>(J->keta2, g);
  }
  return P->jdrift + tot;
}

// European end-of-path jump for NV lockstep slots. All slots share the
// Poisson uniform; slots needing more draws read them from a copy of the
// stream, and the stream advances by what the base path (slot 0) consumed.
template <typename R, int NV, bool TD = false, int KT = 0, typename Gen>
__device__ __forceinline__ void euro_jump(const Job<R> *__restrict__ J,
                                          const VarParams<R> *__restrict__ const *P, R *x, Gen &g,
                                          double *xt = nullptr) {
  if (J->jump_kind == 0) return;
  const double ua = uniform(g);
  int n[NV];
  int nmax = 0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    n[k] = jump_count(P[k]->jmean, P[k]->jexpm, ua);
    nmax = n[k] > nmax ? n[k] : nmax;
  }
  R tot[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = 0;
  if (nmax > 0) {
    Gen t = g;
    if (J->jump_kind == 1) {
      const R Z = normcdfinv<R>(uniform(t));
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (n[k] > 0)
          tot[k] = static_cast<R>(n[k]) * J->jmu + dsqrt<R>(static_cast<R>(n[k])) * J->jdelta * Z;
      if (n[0] > 0) g = t;
    } else {
      for (int i = 0; i < nmax; ++i) {
        const double u = uniform(t);
        const R E = static_cast<R>(-log1p(-uniform(t)));
        const R jsz = (u < static_cast<double>(J->kp)) ? E / J->keta1 : -E / J->keta2;
#pragma unroll
        for (int k = 0; k < NV; ++k)
          if (i < n[k]) tot[k] += jsz;
        if (i + 1 == n[0]) g = t;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if constexpr (TD) {
      if (k == KT || k == KT + 1) {
        xt[k - KT] += P[k]->jdrift_d + static_cast<double>(tot[k]);
        continue;
      }
    }
    x[k] += P[k]->jdrift + tot[k];
  }
}

// ---- rare Giles branches, packed across the wavefront ----------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u)));
}

// Slots whose Poisson inversion hit one of the rare branches (bit k of
// `rare`) are numbered slot-major across the active lanes of the wave
// (ballot + mbcnt), written to this wave's LDS list, inverted by the active
// lanes one item each, and read back by their owners. The wave sits through
// the rare code once per 64 items instead of once per slot. Must be entered
// by every active lane of the wave (a wave-uniform condition).
//
// poisson_fp32_item: one packed item's FP32 first pass on its own intensity,
// with the tail branch and the Temme correction. The single-slot paths (the
// multilevel kernel's steps among them) call it out of line: inlined into
// that kernel's step loop, its code made the compiler turn every job-field
// read of the loop into a vector load (694 -> 1301 ms for the Heston
// up-and-out run, which never takes this path). The lockstep path inlines
// it (5 ms of 371 on the Bates strip).
__device__ static __noinline__ int poisson_fp32_item(double lam, double u) {
  return poisson_first32<true>(lam, u, lam > 4.0 ? giles_w32(u) : 0.0f, nullptr);
}
template <int NV, typename Lam>
__device__ __forceinline__ void giles_rare(const Lam &lams, double U, unsigned rare, int (&Ns)[NV]) {
  constexpr int W64 = 64;
  __shared__ double s_lam[BLOCK / W64][W64];
  __shared__ double s_u[BLOCK / W64][W64];
  __shared__ int s_n[BLOCK / W64][W64];
  const int wave = threadIdx.x / W64;
  const unsigned long long active = __ballot(1);
  const int n_active = __popcll(active);
  const int rank = lanes_below(active);
  int idx[NV];
  int total = 0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const unsigned long long b = __ballot((rare >> k) & 1u);
    idx[k] = total + lanes_below(b);
    total += __popcll(b);
  }
  for (int base = 0; base < total; base += n_active) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int g = idx[k] - base;
      if (((rare >> k) & 1u) && g >= 0 && g < n_active) {
        s_lam[wave][g] = lams(k);
        s_u[wave][g] = U;
      }
    }
    wave_sync();
    if (rank < total - base) {
      // the FP32 first pass with its tail branch on the item's own
      // intensity, then the FP64 method where that is undecided
      const double lam = s_lam[wave][rank], u = s_u[wave][rank];
      int n;
      if constexpr (NV > 1) n = poisson_first32<true>(lam, u, lam > 4.0 ? giles_w32(u) : 0.0f, nullptr);
      else n = poisson_fp32_item(lam, u);
      if (n < 0) n = static_cast<int>(poissinv(lam, u));
      s_n[wave][rank] = n;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int g = idx[k] - base;
      if (((rare >> k) & 1u) && g >= 0 && g < n_active) Ns[k] = s_n[wave][g];
    }
    wave_sync();
  }
}

// Count of a bumped slot (intensity lamk) from slot 0's FP32 first pass
// (intensity lam0 = 1/il0, count n0, margins mg) and the shared normal
// quantile W; -1 when undecided, and the slot then takes the full inversion.
//  - Equal intensities: n0 (same inputs, same arithmetic).
//  - Slot 0 decided by the summation: dF(n)/dlam = -p(n; lam) < 0, so raising
//    lam only endangers F(n0) >= U and lowering it only U > F(n0 - 1). With
//    d = |lamk - lam0| and x = d (1 + n0/lam0) <= 1/2, p(n; lam) <= p(n; lam0)
//    e^x <= p(n; lam0) (1 + 2x) between the intensities, so a margin above
//    d (1 + 2x) p keeps n0. Otherwise the margin moves by the integral of
//    p(n; lam) over the interval, p d (1 +- (n/lam0 - 1) d/2) to second
//    order, with the Taylor remainder at most d^3/6 max|p''| and
//    |p''| <= p e^x ((2n/lam0 + 1)^2 + 4n/lam0^2) for lam >= lam0/2; a moved
//    margin clear of that bound gives n0 or the neighbouring count, whose own
//    probability is bounded below (p(n0 + 1; lamk) >= p(n0 + 1; lam0)(1 - d)
//    when rising, p(n0 - 1; lamk) >= p(n0 - 1; lam0)(1 - (n0 - 1) d/lam0)
//    when falling).
//  - Slot 0 decided by the central asymptotic branch: there |dS/dlam| < 1.82
//    (|W| < 3, lam > 4), so the slot's S lies within 2d of S0; when that
//    interval stays clear of the integers and of the correction band
//    [X, X + 2 Del) (Del scaling as 1/lam), floor(S) is n0. Else the slot
//    reruns the branch in FP32 with the shared W.
// Giles' method returns the exact quantile except where U lies within
// rounding of a CDF value; the first pass's error bounds (kSumTol on the
// margins, s_tol on S) keep those cases out, so a count decided here equals
// the method's.
__device__ __forceinline__ int bumped_count(double lam0, double lamk, int n0, double il0, const PoisMargin &mg) {
  const double dl = fabs(lamk - lam0);
  if (dl == 0.0) return n0;
  if (mg.valid) {
    if (!(lam0 > 0.0 && dl <= 0.5 * lam0)) return -1;
    const double x = dl * (1.0 + static_cast<double>(n0) * il0);
    if (!(x <= 0.5)) return -1;
    const double g = dl * (1.0 + 2.0 * x);
    const bool rise = lamk > lam0;
    const double sl = kSumTol;
    const double up = mg.up, lo = mg.lo, p_hi = mg.p_hi, p_lo = mg.p_lo;
    if (rise) {
      if (!(n0 == 0 || lo > sl)) return -1;
      if (up > g * p_hi + sl) return n0;
    } else {
      if (!(up > sl)) return -1;
      if (n0 == 0 || lo > g * p_lo + sl) return n0;
    }
    const double n = static_cast<double>(rise ? n0 : n0 - 1);
    const double p0 = rise ? p_hi : p_lo;
    const double q = 2.0 * n * il0, c = n * il0 - 1.0;
    const double B = dl * dl * dl * (1.0 / 6.0) * p0 * (1.0 + 2.0 * x) * ((q + 1.0) * (q + 1.0) + 2.0 * q * il0) +
                     sl;
    if (rise) {
      const double m = up - p0 * dl * (1.0 + 0.5 * c * dl);  // F(n0; lamk) - U
      if (m > B) return n0;
      const double n1 = static_cast<double>(n0 + 1);
      if (m < -B && p_hi * lam0 * (1.0 - dl) + (m - B) * n1 > sl * n1) return n0 + 1;
    } else {
      const double m = lo - p0 * dl * (1.0 - 0.5 * c * dl);  // U - F(n0 - 1; lamk)
      if (m > B) return n0;
      if (m < -B && p_lo * (1.0 - n * dl * il0) + (m - B) > sl) return n0 - 1;
    }
    return -1;
  }
  if (mg.asym && dl <= 0.5 * lam0 && lam0 - dl > 4.0) {
    const double f0 = static_cast<double>(mg.S) - static_cast<double>(n0);
    const double e = 2.0 * dl + static_cast<double>(mg.s_tol);
    if (f0 + e < 1.0 && f0 - e >= 2.0 * static_cast<double>(mg.Del) * (1.0 + 2.0 * dl * il0)) return n0;
  }
  if (!(lamk > 4.0 && fabsf(mg.Wf) < 2.999f)) return -1;
  float S, Del, tol;
  return giles_asym32(lamk, mg.Wf, S, Del, tol);
}

// Giles' count for a single slot with the rare branches packed across the
// wavefront (giles_rare); the same count as poissinv. Every active lane of
// the wave must reach the call together.
__device__ __forceinline__ int poisson_count_packed(double lam, double u) {
  const float W = (lam > 4.0 && u > 0.0) ? giles_w32(u) : 0.0f;
  int Ns[1] = {poisson_first32(lam, u, W, nullptr)};
  const bool rare = Ns[0] < 0;
  if (__any(rare)) {
    giles_rare<1>([&](int) { return lam; }, u, rare ? 1u : 0u, Ns);
  }
  return Ns[0];
}

// Poisson counts of NV slots (slot 0 the base path, the others bumped) from
// their shared uniform u1, each equal to what Giles' method returns for that
// slot's intensity:
//  - slot 0 takes the FP32 first pass (normal quantile W of u1 computed
//    once); its bottom-up sum also yields the CDF margins F(N0) - u1 and
//    u1 - F(N0 - 1);
//  - a bumped slot k takes N0 when its intensity equals slot 0's, else its
//    count follows from those margins moved by the intensity change, or from
//    slot 0's asymptotic estimate, or from rerunning the asymptotic branch
//    alone with the same W (bumped_count);
//  - anything else (the first pass's undecided cases and the method's rare
//    branches) is finished by the FP64 method on wave-packed items
//    (giles_rare).
// Every active lane of the wave must reach the call together.
// lams(k): slot k's intensity, evaluated where needed rather than kept in
// registers for all slots at once.
template <int NV, typename Lam>
__device__ __forceinline__ void poisson_counts_lockstep(const Lam &lams, double u1, int (&Ns)[NV]) {
  bool big = false;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    big = big || lams(k) > 4.0;
    Ns[k] = 0;
  }
  if (u1 > 0.0) {  // u1 == 0: every count is 0
    const float W = big ? giles_w32(u1) : 0.0f;
    PoisMargin mg;
    const double lam0 = lams(0);
    const int n0 = poisson_first32(lam0, u1, W, &mg);
    const bool r0 = n0 < 0;
    unsigned pend = r0 ? 1u : 0u;
    Ns[0] = r0 ? 0 : n0;
    const double il0 = 1.0 / lam0;
#pragma unroll
    for (int k = 1; k < NV; ++k) {
      const int n = r0 ? -1 : bumped_count(lam0, lams(k), n0, il0, mg);
      if (n >= 0) Ns[k] = n;
      else pend |= 1u << k;
    }
    if (__any(pend != 0u)) giles_rare<NV>(lams, u1, pend, Ns);
  }
}

// NCI upper tail (u >= 0.999), row N of tail rows `tail` (element type T):
// cubic Hermite in w = -log(1 - u) with exact slopes, before the slot's
// scale factor
template <typename T, typename R>
__device__ __forceinline__ T nci_tail_rows(const Job<R> *__restrict__ J, const T *__restrict__ tail, int N,
                                           double u2) {
  // u2 passes through an opaque copy: the tail is read for u2 >= 0.999 only,
  // and without it the compiler forms this FP64 logarithm (~140 VALU
  // instructions) on every step, ahead of the branches that read it
  double w = u2;
  asm volatile("" : "+v"(w));
  double tp = (-log1p(-w) - J->tail_w0) * J->tail_inv_hw;
  if (tp < 0.0) tp = 0.0;
  int j = static_cast<int>(tp);
  if (j >= J->tail_pts - 1) j = J->tail_pts - 2;
  const T tt = static_cast<T>(tp - static_cast<double>(j)), tt2 = tt * tt, tt3 = tt2 * tt;
  const auto row = gmem(tail + 2 * (static_cast<size_t>(N) * J->tail_pts + j));
  return (T(2) * tt3 - T(3) * tt2 + T(1)) * row[0] + (tt3 - T(2) * tt2 + tt) * row[1] +
         (T(-2) * tt3 + T(3) * tt2) * row[2] + (tt3 - tt2) * row[3];
}

template <typename R>
__device__ __forceinline__ R nci_tail_value(const Job<R> *__restrict__ J, int f, int N, double u2) {
  return nci_tail_rows<R, R>(J, static_cast<const R *>(J->nci_tail[f]), N, u2);
}

// Row N of the FP64 main table at grid interval i, fraction t
__device__ __forceinline__ double nci_row64(const double *__restrict__ tab, int grid, int N, int i, double t) {
  const double t2 = t * t, t3 = t2 * t;
  const auto row = gmem(tab + 2 * (static_cast<size_t>(N) * grid + i));
  return (2.0 * t3 - 3.0 * t2 + 1.0) * row[0] + (t3 - 2.0 * t2 + t) * row[1] + (-2.0 * t3 + 3.0 * t2) * row[2] +
         (t3 - t2) * row[3];
}

// ---- one variance factor step for all NV slots (VolSchemes.hpp) -----------
// TD: slots KT, KT + 1 are stepped in FP64 from vt into vnt (their entries
// of v / vn are not used).
template <typename R, int NV, int NFX, bool TD = false, int KT = 0, typename Gen>
__device__ __forceinline__ void evolve_factor(const Job<R> *__restrict__ J,
                                              const VarParams<R> *__restrict__ const *P, int f,
                                              R (&v)[NV][NFX], R (&vn)[NV][NFX], Gen &g,
                                              double (*vt)[NFX] = nullptr, double (*vnt)[NFX] = nullptr,
                                              int *n_out = nullptr) {
  const int sch = J->scheme[f];
  auto is_t = [](int k) { return TD && (k == KT || k == KT + 1); };
  // a slot's new variance from its (unscaled) table value or exact quantile
  auto put = [&](int k, R y) {
    if (is_t(k)) vnt[k - KT][f] = static_cast<double>(y) * P[k]->Kd[f][2];
    else vn[k][f] = y * P[k]->Kr[f][2];
  };
  auto putd = [&](int k, double q) {
    if (is_t(k)) vnt[k - KT][f] = q * P[k]->Kd[f][2];
    else vn[k][f] = static_cast<R>(q * P[k]->Kd[f][2]);
  };
  if (sch == 1) {  // Ninomiya-Victoir
    const R Z = normcdfinv<R>(uniform(g));
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      if (is_t(k)) {
        const double *K = P[k]->Kd[f];
        const double root = sqrt_nr(K[0] * vt[k - KT][f] + K[1]) + K[2] * static_cast<double>(Z);
        vnt[k - KT][f] = K[0] * (root * root) + K[1];
        continue;
      }
      const R root = dsqrt<R>(P[k]->Kr[f][0] * v[k][f] + P[k]->Kr[f][1]) + P[k]->Kr[f][2] * Z;
      vn[k][f] = P[k]->Kr[f][0] * (root * root) + P[k]->Kr[f][1];
    }
  } else if (sch == 2) {  // NCI: Poisson mixture, tabulated chi2 inverse
    const double u1 = uniform(g);
    const double u2 = uniform(g);
    int Ns[NV];
    if constexpr (NV == 1) {
      Ns[0] = poisson_count_packed(P[0]->Kd[f][0] * static_cast<double>(v[0][f]), u1);
    } else {
      auto lam = [&](int k) {
        return P[k]->Kd[f][0] * (is_t(k) ? vt[k - KT][f] : static_cast<double>(v[k][f]));
      };
      if constexpr (sizeof(R) == 4) {
        // FP32 kernels form each intensity where it is read: five FP64
        // intensities kept live across the inversion pushed the four-wave
        // build's spills from 15 to 29 VGPRs
        poisson_counts_lockstep<NV>(lam, u1, Ns);
      } else {
        // the FP64 kernel (three waves, no spills) runs faster with them
        // formed once
        double lams[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) lams[k] = lam(k);
        poisson_counts_lockstep<NV>([&](int k) { return lams[k]; }, u1, Ns);
      }
    }
    // grid interval and Hermite weights of u2, formed after the counts so
    // they are not live across the Poisson inversion
    const double pos = u2 * J->nci_inv_dx;
    int i = static_cast<int>(pos);
    if (i >= J->nci_grid - 1) i = J->nci_grid - 2;
    const R t = static_cast<R>(pos - static_cast<double>(i));
    const R t2 = t * t, t3 = t2 * t;
    const R h00 = R(2) * t3 - R(3) * t2 + R(1), h10 = t3 - R(2) * t2 + t;
    const R h01 = R(-2) * t3 + R(3) * t2, h11 = t3 - t2;
    const R *__restrict__ tab = static_cast<const R *>(J->nci[f]);
    // the table (chi2 quantiles of d + 2N degrees of freedom) and u2 are
    // common to all slots, so slots sharing slot 0's count share its
    // interpolated quantile and differ only in their scale factor
    // Rows N >= 2 have quantile 0 at the table's u = 0 node but rise
    // steeply just above it (like u^(2/(d+2N))), so interpolating the first
    // grid interval biases the draw low (by up to ~3e-5 of the row mean per
    // step for large N); below nci_u_lo those rows are inverted exactly
    const bool lo = u2 < J->nci_u_lo;
    const bool acc = u2 >= J->nci_u_acc && (u2 < J->nci_u_acc_hi || u2 >= 0.999);
    // row N is read from the table for slot k (else inverted exactly)
    auto in_tab = [&](int N, int k) {
      return (N <= P[k]->nmax[f] && !(lo && N >= 2)) || (N <= J->nci_rows_acc[f] && acc);
    };
    const bool main0 = in_tab(Ns[0], 0) && u2 < 0.999;
    R y0 = R(0);
    // The maturity-bumped slots of FP32 kernels read the FP64 rows (below).
    // There slot 0's row is read from the FP64 table alone -- one gather
    // per step instead of an FP32 and an FP64 one -- and the FP32 slots take
    // its value rounded to FP32.
    double td = 0.0, y0d = 0.0;
    if constexpr (TD) td = pos - static_cast<double>(i);
    if (main0) {
      if constexpr (TD) {
        y0d = nci_row64(J->nci64[f], J->nci_grid, Ns[0], i, td);
        y0 = static_cast<R>(y0d);
      } else {
        const auto row = gmem(tab + 2 * (static_cast<size_t>(Ns[0]) * J->nci_grid + i));
        y0 = h00 * row[0] + h10 * row[1] + h01 * row[2] + h11 * row[3];
      }
    }
    // likewise slot 0's upper-tail value, or its exact quantile when its
    // row is inverted exactly: computed once, taken by the slots on the same
    // row and path (FP32 paths; the FP64 kernel runs faster without the
    // extra live state)
    bool ex0 = false, tl0 = false;
    double q0 = 0.0;
    R yt0 = R(0);
    if constexpr (NV > 1 && sizeof(R) == 4) {
      const bool in0 = in_tab(Ns[0], 0);
      tl0 = in0 && u2 >= 0.999 && J->nci_tail[f] != nullptr;
      ex0 = !(in0 && (u2 < 0.999 || J->nci_tail[f] != nullptr));
      if (tl0) yt0 = nci_tail_value<R>(J, f, Ns[0], u2);
      if (ex0) {
        const double dof = P[0]->Kd[f][1] + 2.0 * Ns[0];
        const double lg = (Ns[0] < J->nci_lg_n) ? gmem(J->nci_lgam[f])[Ns[0]] : lgamma(0.5 * dof);
        q0 = chi2_quantile_lg(dof, u2, lg);
      }
    }
    // exact chi-squared quantile of slot k on row N (slot 0's when shared)
    auto exact_q = [&](int k, int N) {
      if (ex0 && N == Ns[0] && P[k]->Kd[f][1] == P[0]->Kd[f][1]) return q0;
      const double dof = P[k]->Kd[f][1] + 2.0 * N;
      const double lg = (N < J->nci_lg_n) ? gmem(J->nci_lgam[f])[N] : lgamma(0.5 * dof);
      return chi2_quantile_lg(dof, u2, lg);
    };
    // The maturity-bumped slots of FP32 kernels read the FP64 rows, so their
    // variance paths are the FP64 kernel's: a table value rounded to FP32
    // shifts the next step's Poisson intensity, and where the uniform lies
    // near a CDF value the count, which would part those paths from the FP64
    // ones and from each other.
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int N = Ns[k];
      const bool tab_k = in_tab(N, k);
      if constexpr (TD) {
        if (is_t(k)) {
          double y;
          if (tab_k && u2 < 0.999) y = (main0 && N == Ns[0]) ? y0d : nci_row64(J->nci64[f], J->nci_grid, N, i, td);
          else if (tab_k && J->nci_tail[f] != nullptr) y = nci_tail_rows<double>(J, J->nci_tail64[f], N, u2);
          else y = exact_q(k, N);
          vnt[k - KT][f] = y * P[k]->Kd[f][2];
          continue;
        }
      }
      if (main0 && N == Ns[0] && tab_k) {
        put(k, y0);
      } else if (tab_k && u2 < 0.999) {
        const auto row = gmem(tab + 2 * (static_cast<size_t>(N) * J->nci_grid + i));
        put(k, h00 * row[0] + h10 * row[1] + h01 * row[2] + h11 * row[3]);
      } else if (tab_k && u2 >= 0.999 && J->nci_tail[f] != nullptr) {
        put(k, (tl0 && N == Ns[0]) ? yt0 : nci_tail_value<R>(J, f, N, u2));
      } else {
        putd(k, exact_q(k, N));
      }
    }
    if (n_out) *n_out = Ns[0];  // slot 0's Poisson count (control-variate recording)
  } else {  // Exact (Poisson-gamma mixture); variable draws, so never lockstep
    if constexpr (NV == 1) {
      const double lam = P[0]->Kd[f][0] * static_cast<double>(v[0][f]);
      const int N = poisson_count_packed(lam, uniform(g));
      vn[0][f] = gamma_draw<R>(P[0]->Kd[f][1], N, g) * P[0]->Kr[f][2];
      if (n_out) *n_out = N;
    }
  }
}

// Simulate one variant group for the lane's current sample and stage it.
// PD selects path-dependent trackers; European kernels carry no tracker
// state at all.
template <typename R, int NF, int NV, bool PD, int BLK = BLOCK, typename Gen>
__device__ __forceinline__ void sim_group(const Job<R> *__restrict__ J, int gfirst, Gen &g,
                                          R *__restrict__ stage, int tid) {
  constexpr int NFX = NF > 0 ? NF : 1;
  const VarParams<R> *__restrict__ P[NV];
  int vid[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    vid[k] = J->var_list[gfirst + k];
    P[k] = &J->var[vid[k]];
  }
  R x[NV], v[NV][NFX];
  Trk<R> tr[PD ? NV : 1];
  // FP32 Europeans carry the log-return x - log S0 (no tracker compares it
  // with a log-space level) and stage S0 e^x: a log-price near log S0 ~ 4.6
  // would round every path's start by up to 2.4e-7 -- one systematic shift of
  // all terminal spots -- and each step's sum by as much again, where the
  // log-return's rounding scales with its own size. FP64 paths keep the
  // log-price, the reference's arithmetic (ASVJstepper.hpp:134-229).
  constexpr bool LR = log_return_paths<R, PD>();
  const R x0 = LR ? R(0) : static_cast<R>(J->logS0);
  // FP64 state of the maturity-bumped slots (t_slots_fp64)
  constexpr bool TD = t_slots_fp64<R, NV, NFX, PD>();
  constexpr int KT = t_slot0<NFX>();
  double xt[2], vt[2][NFX], vnt[2][NFX];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    x[k] = x0;
#pragma unroll
    for (int f = 0; f < NFX; ++f) v[k][f] = P[k]->v0[f];
    if constexpr (PD) trk_init(J, tr[k], x0, P[k]->dt);
  }
  if constexpr (TD) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      xt[i] = 0.0;
#pragma unroll
      for (int f = 0; f < NFX; ++f) vt[i][f] = P[KT + i]->v0d[f];
    }
  }
  const unsigned steps = static_cast<unsigned>(J->steps);
  if (NF == 0 && !PD) {
    // zero-factor European: one step to maturity (ASVJstepper.hpp:145-161)
    const R G = normcdfinv<R>(uniform(g));
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      if (TD && (k == KT || k == KT + 1)) {
        xt[k - KT] += P[k]->gdrift_nd + P[k]->gvol_nd * static_cast<double>(G);
        continue;
      }
      x[k] += P[k]->gdrift_n + P[k]->gvol_n * G;
    }
    euro_jump<R, NV, TD, KT>(J, P, x, g, xt);
  } else {
    for (unsigned step = 0; step < steps; ++step) {
      const R G = normcdfinv<R>(uniform(g));
      if constexpr (NF == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) x[k] += P[k]->gdrift + P[k]->gvol * G;
      } else {
        R vn[NV][NFX];
        // factors in draw order; written out so the factor index is a
        // compile-time constant and the slot arrays stay in registers
        evolve_factor<R, NV, NFX, TD, KT>(J, P, 0, v, vn, g, vt, vnt);
        if constexpr (NF > 1) evolve_factor<R, NV, NFX, TD, KT>(J, P, 1, v, vn, g, vt, vnt);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          if (TD && (k == KT || k == KT + 1)) {
            const int i = k - KT;
            double a = 0, d = 0;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
              const double *C = P[k]->Cd[f];
              a += C[3] * (vt[i][f] + vnt[i][f]);
              d += C[0] + C[1] * vt[i][f] + C[2] * vnt[i][f];
            }
            xt[i] += d + sqrt_nr(a) * static_cast<double>(G);
#pragma unroll
            for (int f = 0; f < NF; ++f) vt[i][f] = vnt[i][f];
            continue;
          }
          R a = 0, d = 0;
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            a += P[k]->C[f][3] * (v[k][f] + vn[k][f]);
            d += P[k]->C[f][0] + P[k]->C[f][1] * v[k][f] + P[k]->C[f][2] * vn[k][f];
          }
          x[k] += d + dsqrt<R>(a) * G;
#pragma unroll
          for (int f = 0; f < NF; ++f) v[k][f] = vn[k][f];
        }
      }
      if constexpr (PD) {
        if (J->jump_kind != 0) {
          if constexpr (NV == 1) x[0] += log_jump(J, P[0], g);
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) trk_update(J, tr[k], x[k], P[k]->dt);
      }
    }
    if constexpr (!PD && NF > 0) euro_jump<R, NV, TD, KT>(J, P, x, g, xt);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (TD && k == KT) {
      const double S = J->S0 * exp(xt[0]);
      const R hi = static_cast<R>(S);
      stage[(F_SPOT * MAXV + vid[k]) * BLK + tid] = hi;
      stage[(F_SPOT * MAXV + V_RUP) * BLK + tid] = static_cast<R>(S - static_cast<double>(hi));
      continue;
    }
    if (TD && k == KT + 1) {
      stage[(F_SPOT * MAXV + vid[k]) * BLK + tid] = static_cast<R>(expm1f(static_cast<float>(xt[1] - xt[0])));
      continue;
    }
    if constexpr (PD) {
      finalize_stage<R, BLK>(J, tr[k], x[k], P[k]->dt, P[k]->total_time, stage, vid[k], tid);
    } else if constexpr (LR) {
      stage[(F_SPOT * MAXV + vid[k]) * BLK + tid] = static_cast<R>(J->S0) * dexp<R>(x[k]);
    } else {
      stage[(F_SPOT * MAXV + vid[k]) * BLK + tid] = dexp<R>(x[k]);
    }
  }
}

// ---- payoffs (Payoffs.hpp:14-69) and per-variant payoff triples -----------
// Digital payoffs, erfc-smoothed with bandwidth eps in Greek modes; kept out
// of line so the strike loop of vanilla strips stays small.
__device__ static __noinline__ double digital_payoff(int kind, bool smooth, double spot, double diff,
                                                     double eps_scale) {
  const double w = smooth ? 0.5 * erfc(-diff * eps_scale) : (diff > 0.0 ? 1.0 : 0.0);
  return kind == 1 ? w : spot * w;
}

template <typename R>
__device__ __forceinline__ double payoff(const Job<R> *__restrict__ J, bool smooth, double spot, double K) {
  const double diff = J->call ? spot - K : K - spot;
  if (J->payoff == 0) return diff > 0.0 ? diff : 0.0;
  return digital_payoff(J->payoff, smooth, spot, diff, J->eps_scale);
}

// base / up / dn payoffs of staged variant v at sample s for path-dependent
// instruments (Asian.hpp:168-206, Lookback.hpp:229-268, Barrier.hpp:76-110)
template <typename R, int BLK = BLOCK>
__device__ static __noinline__ void eval3_pd(const Job<R> *__restrict__ J, const R *__restrict__ stage, int v,
                                             int s, double K, bool bumps, double &p, double &pu, double &pd) {
  const double S = static_cast<double>(stage[(F_SPOT * MAXV + v) * BLK + s]);
  const double mu = J->mult_up, md = J->mult_dn;
  const bool smooth = J->mode != 0;
  const int trk = J->tracker;
  if (trk != 4) {
    const double A = static_cast<double>(stage[(F_AUX * MAXV + v) * BLK + s]);
    if (J->fixed_strike) {
      p = payoff(J, smooth, A, K);
      if (bumps) {
        pu = payoff(J, smooth, A * mu, K);
        pd = payoff(J, smooth, A * md, K);
      }
    } else {
      p = payoff(J, smooth, S, A);
      if (bumps) {
        pu = payoff(J, smooth, S * mu, A * mu);
        pd = payoff(J, smooth, S * md, A * md);
      }
    }
  } else {
    const int hits = static_cast<int>(stage[(F_FLAGS * MAXV + v) * BLK + s]);
    auto bar = [&](double sp, int idx) {
      const bool hit = (hits >> idx) & 1;
      double reb = 0.0;
      if (J->rebate_timing != 0) {
        reb = J->rebate;
        if (J->rebate_timing == 1 && hit) {
          const double ht = static_cast<double>(stage[((F_HT0 + idx) * MAXV + v) * BLK + s]);
          reb = J->rebate * exp(J->rebate_rate * (J->maturity - ht));
        }
      }
      const double pv = payoff(J, false, sp, K);
      if (J->barrier_in) return hit ? pv : reb;
      return hit ? reb : pv;
    };
    p = bar(S, 0);
    if (bumps) {
      pu = bar(S * mu, 1);
      pd = bar(S * md, 2);
    }
  }
}

// base / up / dn payoffs of staged variant v at sample s (calculate_to_buffer;
// European.hpp:40-74 scales the terminal spot by (S0 +- h)/S0). PD = false:
// the kernel prices Europeans only (no tracker, no out-of-line call).
template <typename R, int BLK = BLOCK, bool PD = true>
__device__ __forceinline__ void eval3(const Job<R> *__restrict__ J, const R *__restrict__ stage, int v,
                                      int s, double K, bool bumps, double &p, double &pu, double &pd) {
  if constexpr (PD) {
    if (J->tracker != 0) {
      eval3_pd<R, BLK>(J, stage, v, s, K, bumps, p, pu, pd);
      return;
    }
  }
  const double S = static_cast<double>(stage[(F_SPOT * MAXV + v) * BLK + s]);
  const bool smooth = J->mode != 0;
  p = payoff(J, smooth, S, K);
  if (bumps) {
    pu = payoff(J, smooth, S * J->mult_up, K);
    pd = payoff(J, smooth, S * J->mult_dn, K);
  }
}

// Strike loop over one staged batch: lane (j, grp) adds strike j of samples
// [grp*nkp, grp*nkp + nkp) (MonteCarloPricer.hpp:264-276 and :312-409
// per-path estimators). The statistics are gathered in passes over the
// staged samples -- price/delta/gamma, then vega+vanna per factor, then
// theta+charm, then rho -- so at most six FP64 sums are live at once; each
// pass adds its batch sums into the lane's own slots in global memory
// (accg[a*BLK + tid]), so the path simulation holds no accumulators.
template <int BLK>
__device__ __forceinline__ void flush(double *__restrict__ accg, int a, int tid, bool first, double x) {
  const auto slot = gmem_w(accg + a * BLK + tid);
  *slot = first ? x : *slot + x;
}

// base / up / dn payoffs of a European at terminal spot S (FP64)
template <typename R>
__device__ __forceinline__ void eval3_spot(const Job<R> *__restrict__ J, double S, double K, double &p, double &pu,
                                           double &pd) {
  const bool smooth = J->mode != 0;
  p = payoff(J, smooth, S, K);
  pu = payoff(J, smooth, S * J->mult_up, K);
  pd = payoff(J, smooth, S * J->mult_dn, K);
}

// End of strike group [s0, s0 + nkp)'s valid samples. A batch's valid
// samples form a prefix of each wavefront's lanes (a lane's sample count
// falls as its stream index rises; a quasi-Monte Carlo wave's points are
// valid up to a lane), and the group's samples are lanes of this lane's own
// wavefront, so its valid samples are [s0, valid_end). Every lane of the
// wavefront must make the call.
__device__ __forceinline__ int valid_end(const int *__restrict__ valid, int tid, int s0, int nkp) {
  const uint64_t vm = __ballot(valid[tid] != 0);
  const uint64_t gm = nkp == 64 ? vm : (vm >> (s0 & 63)) & ((uint64_t{1} << nkp) - 1);
  return s0 + __popcll(gm);
}

// REL: the maturity-bumped variants are staged in FP64 form (t_slots_fp64).
// PD: as in eval3.
template <typename R, int BLK, int UNR, bool REL = false, bool PD = true>
__device__ __forceinline__ void accumulate(const Job<R> *__restrict__ J, const R *__restrict__ stage,
                                           const int *__restrict__ valid, int tid, double *__restrict__ accg,
                                           bool first) {
  // An opaque zero on the lane index keeps the loop-invariant parts of this
  // function (strike, accumulator slot addresses) from being hoisted out of
  // the sample loop, where they would stay live across the whole path
  // simulation: 155 -> 130 VGPRs for the Bates strip kernel, and the
  // four-wave build's spills drop from 26 to 3.
  int zero = 0;
  asm volatile("" : "+s"(zero));
  tid += zero;
  const int nkp = J->nkp;
  const int j = tid & (nkp - 1);
  const int grp = tid / nkp;
  const int mode = J->mode;
  const bool ess = mode != 0, full = mode == 2;
  const bool live = j < J->nk;
  const double K = live ? gmem(J->strikes)[j] : 0.0;
  // samples are checked one by one here: counting them once per batch
  // (valid_end) took the batched Double-Heston kernel from 127 to 134 VGPRs,
  // one wave per SIMD fewer
  const int s0 = grp * nkp, s1 = grp * nkp + nkp;
  {
    double sp = 0, sp2 = 0, sd = 0, sd2 = 0, sg = 0, sg2 = 0;
    if (live) {
#pragma unroll UNR
      for (int s = s0; s < s1; ++s) {
        if (!valid[s]) continue;
        double p, pu = 0, pd = 0;
        eval3<R, BLK, PD>(J, stage, V_BASE, s, K, ess, p, pu, pd);
        sp += p;
        sp2 += p * p;
        if (ess) {
          const double d = (pu - pd) * J->inv2h;
          const double gm = (pu - p * 2.0 + pd) * J->invh2;
          sd += d;
          sd2 += d * d;
          sg += gm;
          sg2 += gm * gm;
        }
      }
    }
    flush<BLK>(accg, 2 * ST_PRICE, tid, first, sp);
    flush<BLK>(accg, 2 * ST_PRICE + 1, tid, first, sp2);
    if (ess) {
      flush<BLK>(accg, 2 * ST_DELTA, tid, first, sd);
      flush<BLK>(accg, 2 * ST_DELTA + 1, tid, first, sd2);
      flush<BLK>(accg, 2 * ST_GAMMA, tid, first, sg);
      flush<BLK>(accg, 2 * ST_GAMMA + 1, tid, first, sg2);
    }
  }
  if (!full) return;
#pragma unroll 1
  for (int f = 0; f < J->nvega; ++f) {
    double sv = 0, sv2 = 0, sn = 0, sn2 = 0;
    if (live) {
#pragma unroll UNR
      for (int s = s0; s < s1; ++s) {
        if (!valid[s]) continue;
        double ap, au, ad, bp, bu, bd;
        eval3<R, BLK, PD>(J, stage, V_V0UP + 2 * f, s, K, true, ap, au, ad);
        eval3<R, BLK, PD>(J, stage, V_V0DN + 2 * f, s, K, true, bp, bu, bd);
        const double vg = (ap - bp) * J->inv2dv * J->cv[f];
        const double vn = ((au - ad) - (bu - bd)) * (J->inv2dv * J->inv2h) * J->cv[f];
        sv += vg;
        sv2 += vg * vg;
        sn += vn;
        sn2 += vn * vn;
      }
    }
    flush<BLK>(accg, 2 * (ST_VEGA0 + f), tid, first, sv);
    flush<BLK>(accg, 2 * (ST_VEGA0 + f) + 1, tid, first, sv2);
    flush<BLK>(accg, 2 * (ST_VANNA0 + f), tid, first, sn);
    flush<BLK>(accg, 2 * (ST_VANNA0 + f) + 1, tid, first, sn2);
  }
  if (J->nvega < 2) {
    flush<BLK>(accg, 2 * ST_VEGA1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VEGA1 + 1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VANNA1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VANNA1 + 1, tid, first, 0.0);
  }
  {
    double st = 0, st2 = 0, sc = 0, sc2 = 0;
    if (live) {
#pragma unroll UNR
      for (int s = s0; s < s1; ++s) {
        if (!valid[s]) continue;
        double ap, au, ad, bp, bu, bd;
        if constexpr (REL) {
          const double Su = static_cast<double>(stage[(F_SPOT * MAXV + V_TUP) * BLK + s]) +
                            static_cast<double>(stage[(F_SPOT * MAXV + V_RUP) * BLK + s]);
          const double Sd = Su + Su * static_cast<double>(stage[(F_SPOT * MAXV + V_TDN) * BLK + s]);
          eval3_spot(J, Su, K, ap, au, ad);
          eval3_spot(J, Sd, K, bp, bu, bd);
        } else {
          eval3<R, BLK, PD>(J, stage, V_TUP, s, K, true, ap, au, ad);
          eval3<R, BLK, PD>(J, stage, V_TDN, s, K, true, bp, bu, bd);
        }
        const double th = (ap * J->dfTu - bp * J->dfTd) * J->inv2dT;
        const double ch = ((au - ad) * J->dfTu - (bu - bd) * J->dfTd) * (J->inv2dT * J->inv2h);
        st += th;
        st2 += th * th;
        sc += ch;
        sc2 += ch * ch;
      }
    }
    flush<BLK>(accg, 2 * ST_THETA, tid, first, st);
    flush<BLK>(accg, 2 * ST_THETA + 1, tid, first, st2);
    flush<BLK>(accg, 2 * ST_CHARM, tid, first, sc);
    flush<BLK>(accg, 2 * ST_CHARM + 1, tid, first, sc2);
  }
  {
    double sr = 0, sr2 = 0;
    if (live) {
#pragma unroll UNR
      for (int s = s0; s < s1; ++s) {
        if (!valid[s]) continue;
        double ap, bp, dum0, dum1;
        if (J->derive_r) {
          // European: the rate-bumped paths are the base path shifted by
          // +-r_bump*T in log space (only C0 of the log step depends on r)
          const double S = static_cast<double>(stage[(F_SPOT * MAXV + V_BASE) * BLK + s]);
          ap = payoff(J, mode != 0, S * J->erb_up, K);
          bp = payoff(J, mode != 0, S * J->erb_dn, K);
        } else {
          eval3<R, BLK, PD>(J, stage, V_RUP, s, K, false, ap, dum0, dum1);
          eval3<R, BLK, PD>(J, stage, V_RDN, s, K, false, bp, dum0, dum1);
        }
        const double rh = (ap * J->dfru - bp * J->dfrd) * J->inv2dr;
        sr += rh;
        sr2 += rh * rh;
      }
    }
    flush<BLK>(accg, 2 * ST_RHO, tid, first, sr);
    flush<BLK>(accg, 2 * ST_RHO + 1, tid, first, sr2);
  }
}

// accumulate() for European vanilla calls and puts, the strips' common case,
// with every job field it reads loaded once per batch and the payoff written
// out. With sg = +1 (call) or -1 (put), sg (S m - K) = fma(sg S, m, -sg K)
// exactly (negation is exact and rounding symmetric), so each payoff is one
// FMA (or subtraction) and a max, the same number payoff() returns.
template <typename R, int BLK, int UNR, bool REL>
__device__ __forceinline__ void accumulate_vanilla(const Job<R> *__restrict__ J, const R *__restrict__ stage,
                                                   const int *__restrict__ valid, int tid,
                                                   double *__restrict__ accg, bool first) {
  int zero = 0;  // as in accumulate: keeps the strike and slot addresses in the batch
  asm volatile("" : "+s"(zero));
  tid += zero;
  const int nkp = J->nkp;
  const int j = tid & (nkp - 1);
  const int s0 = (tid / nkp) * nkp, s1 = valid_end(valid, tid, s0, nkp);
  const int mode = J->mode;
  const bool ess = mode != 0, full = mode == 2;
  const bool live = j < J->nk;
  const double sg = J->call ? 1.0 : -1.0;
  const double Ks = live ? sg * gmem(J->strikes)[j] : 0.0;
  const int n = live ? s1 - s0 : 0;
  const double mu = J->mult_up, md = J->mult_dn;
  auto spot = [&](int v, int s) { return sg * static_cast<double>(stage[(F_SPOT * MAXV + v) * BLK + s]); };
  auto pay = [&](double Ss) {
    const double d = Ss - Ks;
    return d > 0.0 ? d : 0.0;
  };
  auto pay_m = [&](double Ss, double m) {
    const double d = fma(Ss, m, -Ks);
    return d > 0.0 ? d : 0.0;
  };
  if (!ess) {
    double sp = 0, sp2 = 0;
#pragma unroll UNR
    for (int i = 0; i < n; ++i) {
      const double p = pay(spot(V_BASE, s0 + i));
      sp += p;
      sp2 += p * p;
    }
    flush<BLK>(accg, 2 * ST_PRICE, tid, first, sp);
    flush<BLK>(accg, 2 * ST_PRICE + 1, tid, first, sp2);
    return;
  }
  {
    const double inv2h = J->inv2h, invh2 = J->invh2;
    double sp = 0, sp2 = 0, sd = 0, sd2 = 0, sg2 = 0, sgm = 0;
#pragma unroll UNR
    for (int i = 0; i < n; ++i) {
      const double Ss = spot(V_BASE, s0 + i);
      const double p = pay(Ss), pu = pay_m(Ss, mu), pd = pay_m(Ss, md);
      sp += p;
      sp2 += p * p;
      const double d = (pu - pd) * inv2h;
      const double gm = (pu - p * 2.0 + pd) * invh2;
      sd += d;
      sd2 += d * d;
      sgm += gm;
      sg2 += gm * gm;
    }
    flush<BLK>(accg, 2 * ST_PRICE, tid, first, sp);
    flush<BLK>(accg, 2 * ST_PRICE + 1, tid, first, sp2);
    flush<BLK>(accg, 2 * ST_DELTA, tid, first, sd);
    flush<BLK>(accg, 2 * ST_DELTA + 1, tid, first, sd2);
    flush<BLK>(accg, 2 * ST_GAMMA, tid, first, sgm);
    flush<BLK>(accg, 2 * ST_GAMMA + 1, tid, first, sg2);
  }
  if (!full) return;
#pragma unroll 1
  for (int f = 0; f < J->nvega; ++f) {
    const double cvg = J->inv2dv, cvn = J->inv2dv * J->inv2h, cf = J->cv[f];
    double sv = 0, sv2 = 0, sn = 0, sn2 = 0;
#pragma unroll UNR
    for (int i = 0; i < n; ++i) {
      const double A = spot(V_V0UP + 2 * f, s0 + i), B = spot(V_V0DN + 2 * f, s0 + i);
      const double ap = pay(A), au = pay_m(A, mu), ad = pay_m(A, md);
      const double bp = pay(B), bu = pay_m(B, mu), bd = pay_m(B, md);
      const double vg = (ap - bp) * cvg * cf;
      const double vn = ((au - ad) - (bu - bd)) * cvn * cf;
      sv += vg;
      sv2 += vg * vg;
      sn += vn;
      sn2 += vn * vn;
    }
    flush<BLK>(accg, 2 * (ST_VEGA0 + f), tid, first, sv);
    flush<BLK>(accg, 2 * (ST_VEGA0 + f) + 1, tid, first, sv2);
    flush<BLK>(accg, 2 * (ST_VANNA0 + f), tid, first, sn);
    flush<BLK>(accg, 2 * (ST_VANNA0 + f) + 1, tid, first, sn2);
  }
  if (J->nvega < 2) {
    flush<BLK>(accg, 2 * ST_VEGA1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VEGA1 + 1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VANNA1, tid, first, 0.0);
    flush<BLK>(accg, 2 * ST_VANNA1 + 1, tid, first, 0.0);
  }
  {
    const double dfTu = J->dfTu, dfTd = J->dfTd, inv2dT = J->inv2dT, cch = J->inv2dT * J->inv2h;
    double st = 0, st2 = 0, sc = 0, sc2 = 0;
#pragma unroll UNR
    for (int i = 0; i < n; ++i) {
      const int s = s0 + i;
      double A, B;
      if constexpr (REL) {
        const double Su = static_cast<double>(stage[(F_SPOT * MAXV + V_TUP) * BLK + s]) +
                          static_cast<double>(stage[(F_SPOT * MAXV + V_RUP) * BLK + s]);
        A = sg * Su;
        B = sg * (Su + Su * static_cast<double>(stage[(F_SPOT * MAXV + V_TDN) * BLK + s]));
      } else {
        A = spot(V_TUP, s);
        B = spot(V_TDN, s);
      }
      const double ap = pay(A), au = pay_m(A, mu), ad = pay_m(A, md);
      const double bp = pay(B), bu = pay_m(B, mu), bd = pay_m(B, md);
      const double th = (ap * dfTu - bp * dfTd) * inv2dT;
      const double ch = ((au - ad) * dfTu - (bu - bd) * dfTd) * cch;
      st += th;
      st2 += th * th;
      sc += ch;
      sc2 += ch * ch;
    }
    flush<BLK>(accg, 2 * ST_THETA, tid, first, st);
    flush<BLK>(accg, 2 * ST_THETA + 1, tid, first, st2);
    flush<BLK>(accg, 2 * ST_CHARM, tid, first, sc);
    flush<BLK>(accg, 2 * ST_CHARM + 1, tid, first, sc2);
  }
  {
    const double dfru = J->dfru, dfrd = J->dfrd, inv2dr = J->inv2dr;
    const double eu = J->erb_up, ed = J->erb_dn;
    const bool derive = J->derive_r != 0;
    double sr = 0, sr2 = 0;
#pragma unroll UNR
    for (int i = 0; i < n; ++i) {
      const int s = s0 + i;
      double ap, bp;
      if (derive) {
        // the rate-bumped paths are the base path shifted by +-r_bump*T in
        // log space (only C0 of the log step depends on r)
        const double Ss = spot(V_BASE, s);
        ap = pay_m(Ss, eu);
        bp = pay_m(Ss, ed);
      } else {
        ap = pay(spot(V_RUP, s));
        bp = pay(spot(V_RDN, s));
      }
      const double rh = (ap * dfru - bp * dfrd) * inv2dr;
      sr += rh;
      sr2 += rh * rh;
    }
    flush<BLK>(accg, 2 * ST_RHO, tid, first, sr);
    flush<BLK>(accg, 2 * ST_RHO + 1, tid, first, sr2);
  }
}

// Per-unit partial sums of job J from its lanes' accumulator slots: the
// sample groups of each strike by wave butterfly, then the waves in order;
// one FP64 partial per unit, accumulator and strike.
template <typename R, int BLK>
__device__ __forceinline__ void unit_partials(const Job<R> *__restrict__ J, const double *__restrict__ accg,
                                              int blk, bool any, double *__restrict__ s_red, int tid) {
  const int nkp = J->nkp;
  const int lane = tid & 63, wave = tid >> 6;
  const int mode = J->mode;
  const int na = mode == 2 ? 2 * NSTAT : (mode == 1 ? 2 * (ST_GAMMA + 1) : 2);
  double *out = J->partial + static_cast<size_t>(J->first_slot + blk) * (2 * NSTAT) * J->nk;
  for (int a = 0; a < 2 * NSTAT; ++a) {
    double x = (a < na && any) ? gmem(accg)[a * BLK + tid] : 0.0;
    for (int off = nkp; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
    s_red[tid] = x;
    __syncthreads();
    if (wave == 0 && lane < J->nk) {
      double t = 0.0;
      for (int w = 0; w < BLK / 64; ++w) t += s_red[w * 64 + lane];
      gmem_w(out)[a * J->nk + lane] = t;
    }
    __syncthreads();
  }
}

// Persistent work-queue kernel. `counter` must be zero at launch; `accg`
// holds 2*NSTAT*BLK doubles per work unit of BLK streams; UNR unrolls the
// strike loop over staged samples.
//
// QMC (randomized quasi-Monte Carlo): lane s of the job (wave w = s / 64,
// lane l = s % 64) prices at iteration it the point c = (gray(h) << 6) | l of
// the digitally shifted Sobol' sequence, h = w n_it + it, valid while
// 64 h + l < M; the paths draw the point's coordinates in order (Qrng). The
// wave keeps, for every coordinate d < qmc_dims, hi[d] = shift[d] ^ the XOR
// of V[k][d] over the set bits k >= 6 of c in LDS: built at the unit's
// first point, then one XOR per coordinate per iteration (gray(h) and
// gray(h - 1) differ in bit ctz(h) only). The lane's lower bits come from
// qmc_lo[d][l], the XOR of V[k][d] over the set bits k < 6 of l.
template <typename R, int NF, int NV, bool PD, int BLK, int UNR, bool QMC = false>
__device__ __forceinline__ void mc_paths_body(const Job<R> *__restrict__ jobs, const Unit *__restrict__ units,
                                              int n_units, unsigned *__restrict__ counter,
                                              double *__restrict__ accg_all) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_unit;
  __shared__ int s_valid[BLK];
  __shared__ double s_red[BLK];
  __shared__ uint32_t s_qhi[QMC ? (BLK / 64) * QMC_DMAX : 1];
  R *stage = reinterpret_cast<R *>(smem);
  const int tid = threadIdx.x;

  for (;;) {
    if (tid == 0) s_unit = static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();
    if (u >= n_units) break;
    const int job = __builtin_amdgcn_readfirstlane(units[u].job);
    const int blk = __builtin_amdgcn_readfirstlane(units[u].blk);
    const Job<R> *__restrict__ J = jobs + job;
    double *__restrict__ accg = accg_all + static_cast<size_t>(u) * (2 * NSTAT * BLK);

    const uint64_t S = J->S, M = J->M;
    const uint64_t s = static_cast<uint64_t>(blk) * BLK + tid;
    const bool active = s < S;
    const uint64_t q = M / S, rem = M % S;
    uint64_t count = 0, n_it = 0, h0 = 0;
    Rng g;
    uint32_t *qhi = s_qhi + (QMC ? (tid >> 6) * QMC_DMAX : 0);
    if constexpr (!QMC) {
      if (active) {
        g.s0 = gmem(J->streams)[4 * s + 0];
        g.s1 = gmem(J->streams)[4 * s + 1];
        g.s2 = gmem(J->streams)[4 * s + 2];
        g.s3 = gmem(J->streams)[4 * s + 3];
      } else {
        g.s0 = g.s1 = g.s2 = g.s3 = 0;
      }
      count = active ? q + (s < rem ? 1 : 0) : 0;
      n_it = q + (static_cast<uint64_t>(blk) * BLK < rem ? 1 : 0);
    } else {
      n_it = q + (rem ? 1 : 0);
      h0 = (s >> 6) * n_it;
      qmc_hi_init(qhi, J->qmc_V, J->qmc_shift, J->qmc_dims, h0, tid & 63);
      wave_sync();
    }

    for (uint64_t it = 0; it < n_it; ++it) {
      bool valid;
      if constexpr (QMC) {
        const uint64_t h = h0 + it;
        if (it > 0) {
          qmc_hi_step(qhi, J->qmc_V, J->qmc_dims, h, tid & 63);
          wave_sync();
        }
        valid = active && 64 * h + static_cast<uint64_t>(tid & 63) < M;
      } else {
        valid = it < count;
      }
      if (valid) {
        if constexpr (QMC) {
          Qrng qg{qhi, J->qmc_lo + (tid & 63), 0};
          if constexpr (NV > 1) {
            sim_group<R, NF, NV, PD, BLK>(J, 0, qg, stage, tid);
          } else {
            const Qrng snap = qg;
            sim_group<R, NF, 1, PD, BLK>(J, 0, qg, stage, tid);
            for (int gi = 1; gi < J->n_groups; ++gi) {
              Qrng gg = snap;
              sim_group<R, NF, 1, PD, BLK>(J, J->group_first[gi], gg, stage, tid);
            }
          }
        } else if constexpr (NV > 1) {
          // lockstep: one group, the stream advances as the base path does
          sim_group<R, NF, NV, PD, BLK>(J, 0, g, stage, tid);
        } else {
          // replay: every variant restarts from the sample's stream state
          const Rng snap = g;
          sim_group<R, NF, 1, PD, BLK>(J, 0, g, stage, tid);
          for (int gi = 1; gi < J->n_groups; ++gi) {
            Rng gg = snap;
            sim_group<R, NF, 1, PD, BLK>(J, J->group_first[gi], gg, stage, tid);
          }
        }
      }
      s_valid[tid] = valid ? 1 : 0;
      __syncthreads();
      constexpr bool REL = t_slots_fp64<R, NV, (NF > 0 ? NF : 1), PD>();
      bool done = false;
      if constexpr (!PD) {
        if (J->payoff == 0) {
          accumulate_vanilla<R, BLK, UNR, REL>(J, stage, s_valid, tid, accg, it == 0);
          done = true;
        }
      }
      if (!done) accumulate<R, BLK, UNR, REL, PD>(J, stage, s_valid, tid, accg, it == 0);
      if (J->view)
        accumulate<R, BLK, UNR>(jobs + (J->view - 1), stage, s_valid, tid,
                                J->view_accg + static_cast<size_t>(u) * (2 * NSTAT * BLK), it == 0);
      __syncthreads();
    }

    unit_partials<R, BLK>(J, accg, blk, n_it > 0, s_red, tid);
    if (J->view)
      unit_partials<R, BLK>(jobs + (J->view - 1), J->view_accg + static_cast<size_t>(u) * (2 * NSTAT * BLK), blk,
                            n_it > 0, s_red, tid);
  }
}

template <typename R, int NF, int NV, bool PD, int BLK = BLOCK, int UNR = 4>
__global__ void __launch_bounds__(BLK) mc_paths_kernel(const Job<R> *__restrict__ jobs,
                                                         const Unit *__restrict__ units, int n_units,
                                                         unsigned *__restrict__ counter,
                                                         double *__restrict__ accg_all) {
  mc_paths_body<R, NF, NV, PD, BLK, UNR>(jobs, units, n_units, counter, accg_all);
}

// The same kernel asked to fit four waves per SIMD (at most 128 VGPRs): the
// compiler spills what does not fit, and the extra wave hides more of the
// table gathers' latency; the autotuner weighs the two.
template <typename R, int NF, int NV, bool PD, int BLK, int UNR>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4)))
mc_paths_kernel_occ4(const Job<R> *__restrict__ jobs, const Unit *__restrict__ units, int n_units,
                     unsigned *__restrict__ counter, double *__restrict__ accg_all) {
  mc_paths_body<R, NF, NV, PD, BLK, UNR>(jobs, units, n_units, counter, accg_all);
}

// The same kernel drawing randomized quasi-Monte Carlo points (one launch
// shape: 256-lane blocks, strike loop unrolled by 4).
template <typename R, int NF, int NV, bool PD>
__global__ void __launch_bounds__(BLOCK) mc_paths_kernel_qmc(const Job<R> *__restrict__ jobs,
                                                             const Unit *__restrict__ units, int n_units,
                                                             unsigned *__restrict__ counter,
                                                             double *__restrict__ accg_all) {
  mc_paths_body<R, NF, NV, PD, BLOCK, 4, true>(jobs, units, n_units, counter, accg_all);
}

}  // namespace myql
