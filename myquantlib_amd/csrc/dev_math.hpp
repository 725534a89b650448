// Device-side random numbers and special functions for the path kernels
// (gfx950, 64-wide wavefronts). One xoshiro256++ stream lives in four 64-bit
// registers of a lane; every transform below consumes draws exactly as the
// reference's Boost distributions do (MonteCarloPricer.hpp:86,
// ASVJstepper.hpp:137, JumpPolicies.hpp:38-39), so the number of draws per
// path -- and hence the stream position of every later path -- matches the
// FP64 oracle bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace myql {

// A pointer into device (global) memory read out of a parameter block is
// generic to the compiler, which then loads through it with flat
// instructions. Those also count against the scalar loads' wait counter, so
// every later wait for a job field would wait for a table gather in flight
// too; reads through gmem(p) are global loads, which do not.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
  return (const __attribute__((address_space(1))) T *)(p);
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem_w(T *p) {
  return (__attribute__((address_space(1))) T *)(p);
}

struct Rng {
  uint64_t s0, s1, s2, s3;
};

__device__ __forceinline__ uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256++: out = rotl(s0 + s3, 23) + s0 (Blackman & Vigna).
__device__ __forceinline__ uint64_t next_u64(Rng &g) {
  const uint64_t out = rotl(g.s0 + g.s3, 23) + g.s0;
  const uint64_t t = g.s1 << 17;
  g.s2 ^= g.s0;
  g.s3 ^= g.s1;
  g.s1 ^= g.s2;
  g.s0 ^= g.s3;
  g.s2 ^= t;
  g.s3 = rotl(g.s3, 45);
  return out;
}

// advance the state only (jump-ahead inner loop)
__device__ __forceinline__ void step_state(Rng &g) {
  const uint64_t t = g.s1 << 17;
  g.s2 ^= g.s0;
  g.s3 ^= g.s1;
  g.s1 ^= g.s2;
  g.s0 ^= g.s3;
  g.s2 ^= t;
  g.s3 = rotl(g.s3, 45);
}

// uniform_real_distribution<double>(0,1): u = double(x) / 2^64 rounded to
// nearest, redrawn while u == 1 (probability 2^-54).
__device__ __forceinline__ double uniform(Rng &g) {
  for (;;) {
    const double u = static_cast<double>(next_u64(g)) * 0x1p-64;
    if (u < 1.0) return u;
  }
}

// Randomized quasi-Monte Carlo draws (Sobol' points, digitally shifted):
// the d-th uniform of a path is coordinate d of the path's point,
// hi[d] ^ lo[64 d], where hi holds the wavefront's shifted upper-bit part of
// every coordinate (in LDS, updated once per wave step) and lo points at
// the lane's column of the lower-6-bit table (see mc_paths_body); the
// copies a replay takes restart at the same coordinate, as a copied stream
// state would. u = (x + 1/2) 2^-32 lies in (0, 1).
constexpr int QMC_DMAX = 1024;  // coordinates per quasi-Monte Carlo point
struct Qrng {
  const uint32_t *hi;
  const uint32_t *lo;
  int d;
};

// next coordinate as a 32-bit integer
__device__ __forceinline__ uint32_t qmc_next(Qrng &q) {
  // the host sizes the coordinates for the paths' draws; the clamp only
  // keeps every read inside the tables
  const int d = q.d < QMC_DMAX ? q.d : QMC_DMAX - 1;
  const uint32_t x = q.hi[d] ^ q.lo[64 * d];
  ++q.d;
  return x;
}

__device__ __forceinline__ double uniform(Qrng &q) { return (static_cast<double>(qmc_next(q)) + 0.5) * 0x1p-32; }

// The wavefront's upper-bit part of every coordinate (hi[d], d < dims) for
// wave step h: shift[d] ^ the XOR of V[k][d] over the set bits k >= 6 of
// gray(h) << 6. Each lane fills the coordinates d = lane (mod 64).
__device__ __forceinline__ void qmc_hi_init(uint32_t *hi, const uint32_t *__restrict__ V,
                                            const uint32_t *__restrict__ shift, int dims, uint64_t h, int lane) {
  const uint64_t gc = h ^ (h >> 1);
  for (int d = lane; d < dims; d += 64) {
    uint32_t x = shift[d];
    uint64_t b = gc;
    for (int k = 6; b != 0 && k < 32; ++k, b >>= 1)
      if (b & 1) x ^= V[k * QMC_DMAX + d];
    hi[d] = x;
  }
}

// From wave step h - 1 to h (h >= 1): gray(h) differs from gray(h - 1) in
// bit ctz(h) only.
__device__ __forceinline__ void qmc_hi_step(uint32_t *hi, const uint32_t *__restrict__ V, int dims, uint64_t h,
                                            int lane) {
  // points stay below 2^31 (host-checked), so k <= 30; the clamp only keeps
  // the read inside the table
  const int k = min(6 + __builtin_ctzll(h), 31);
  for (int d = lane; d < dims; d += 64) hi[d] ^= V[k * QMC_DMAX + d];
}

__device__ __forceinline__ uint64_t splitmix(uint64_t &x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ Rng seed_rng(uint64_t seed) {
  Rng g;
  uint64_t x = seed;
  g.s0 = splitmix(x);
  g.s1 = splitmix(x);
  g.s2 = splitmix(x);
  g.s3 = splitmix(x);
  return g;
}

// ---------------------------------------------------------------------------
// Acklam's inverse normal CDF (reference Numerics.hpp:12-47). The branch is
// always decided on the FP64 uniform so FP32 and FP64 paths take the same
// region for the same draw; the FP32 form evaluates the rational functions in
// single precision with hardware log2/sqrt/rcp.
// ---------------------------------------------------------------------------
#define MYQL_ACKLAM_A0 -3.969683028665376e+01
#define MYQL_ACKLAM_A1 2.209460984245205e+02
#define MYQL_ACKLAM_A2 -2.759285104469687e+02
#define MYQL_ACKLAM_A3 1.383577518672690e+02
#define MYQL_ACKLAM_A4 -3.066479806614716e+01
#define MYQL_ACKLAM_A5 2.506628277459239e+00
#define MYQL_ACKLAM_B0 -5.447609879822406e+01
#define MYQL_ACKLAM_B1 1.615858368580409e+02
#define MYQL_ACKLAM_B2 -1.556989798598866e+02
#define MYQL_ACKLAM_B3 6.680131188771972e+01
#define MYQL_ACKLAM_B4 -1.328068155288572e+01
#define MYQL_ACKLAM_C0 -7.784894002430293e-03
#define MYQL_ACKLAM_C1 -3.223964580411365e-01
#define MYQL_ACKLAM_C2 -2.400758277161838e+00
#define MYQL_ACKLAM_C3 -2.549732539343734e+00
#define MYQL_ACKLAM_C4 4.374664141464968e+00
#define MYQL_ACKLAM_C5 2.938163982698783e+00
#define MYQL_ACKLAM_D0 7.784695709041462e-03
#define MYQL_ACKLAM_D1 3.224671290700398e-01
#define MYQL_ACKLAM_D2 2.445134137142996e+00
#define MYQL_ACKLAM_D3 3.754408661907416e+00

__device__ __forceinline__ double normcdfinv_d(double p) {
  if (p > 0.02425 && p < 1.0 - 0.02425) {
    const double q = p - 0.5, r = q * q;
    const double num =
        (((((MYQL_ACKLAM_A0 * r + MYQL_ACKLAM_A1) * r + MYQL_ACKLAM_A2) * r + MYQL_ACKLAM_A3) * r +
          MYQL_ACKLAM_A4) * r + MYQL_ACKLAM_A5) * q;
    const double den =
        ((((MYQL_ACKLAM_B0 * r + MYQL_ACKLAM_B1) * r + MYQL_ACKLAM_B2) * r + MYQL_ACKLAM_B3) * r +
         MYQL_ACKLAM_B4) * r + 1.0;
    return num / den;
  }
  const bool lower = p <= 0.02425;
  const double q = lower ? p : 1.0 - p;
  const double r = sqrt(-2.0 * log(q));
  const double x =
      (((((MYQL_ACKLAM_C0 * r + MYQL_ACKLAM_C1) * r + MYQL_ACKLAM_C2) * r + MYQL_ACKLAM_C3) * r +
        MYQL_ACKLAM_C4) * r + MYQL_ACKLAM_C5) /
      ((((MYQL_ACKLAM_D0 * r + MYQL_ACKLAM_D1) * r + MYQL_ACKLAM_D2) * r + MYQL_ACKLAM_D3) * r + 1.0);
  return lower ? x : -x;
}

__device__ __forceinline__ float normcdfinv_f(double p) {
  if (p > 0.02425 && p < 1.0 - 0.02425) {
    // The central rational cancels heavily near the region's edges (terms
    // of size ~7 sum to ~0.02), so it is accumulated in FP64 -- the same
    // VALU rate as FP32 FMA on gfx950 -- and only the quotient is FP32.
    const double q = p - 0.5, r = q * q;
    const double num =
        (((((MYQL_ACKLAM_A0 * r + MYQL_ACKLAM_A1) * r + MYQL_ACKLAM_A2) * r + MYQL_ACKLAM_A3) * r +
          MYQL_ACKLAM_A4) * r + MYQL_ACKLAM_A5) * q;
    const double den =
        ((((MYQL_ACKLAM_B0 * r + MYQL_ACKLAM_B1) * r + MYQL_ACKLAM_B2) * r + MYQL_ACKLAM_B3) * r +
         MYQL_ACKLAM_B4) * r + 1.0;
    return static_cast<float>(num) * __builtin_amdgcn_rcpf(static_cast<float>(den));
  }
  const bool lower = p <= 0.02425;
  // tail argument formed in FP64 so 1 - p keeps its precision near 1
  const float q = static_cast<float>(lower ? p : 1.0 - p);
  const float r = __builtin_amdgcn_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(q));
  const float x =
      (((((float(MYQL_ACKLAM_C0) * r + float(MYQL_ACKLAM_C1)) * r + float(MYQL_ACKLAM_C2)) * r +
         float(MYQL_ACKLAM_C3)) * r + float(MYQL_ACKLAM_C4)) * r + float(MYQL_ACKLAM_C5)) *
      __builtin_amdgcn_rcpf(
          (((float(MYQL_ACKLAM_D0) * r + float(MYQL_ACKLAM_D1)) * r + float(MYQL_ACKLAM_D2)) * r +
           float(MYQL_ACKLAM_D3)) * r + 1.0f);
  return lower ? x : -x;
}

template <typename R>
__device__ __forceinline__ R normcdfinv(double p);
template <>
__device__ __forceinline__ double normcdfinv<double>(double p) { return normcdfinv_d(p); }
template <>
__device__ __forceinline__ float normcdfinv<float>(double p) { return normcdfinv_f(p); }

// Coefficients of the Temme correction in Giles' inverse Poisson CDF, read
// through scalar loads from constant memory.
__constant__ static const double kTemmeB1[13] = {
    8.0995211567045583e-16,  -5.1391118342426808e-16, -1.3923887224181616e-13, 1.1004392031956284e-13,
    2.4361948020667402e-11,  -2.5514193994946487e-11, -4.3820360184533521e-09, 6.7078535434015332e-09,
    8.2967113409530833e-07,  -2.1854485106799979e-06, -0.00017875514403292177, 0.0011574074074074078,
    0.083333333333333329};
__constant__ static const double kTemmeB0[13] = {
    -1.9752288294349411e-15, 2.8534893807047458e-14,  3.3717632624009806e-13,  -5.0276692801141763e-12,
    -5.8307721325504166e-11, 9.1476995822367933e-10,  1.0261809784240299e-08,  -1.7665952736826086e-07,
    -1.8540622107151585e-06, 3.9192631785224383e-05,  0.00035273368606701921,  -0.014814814814814815,
    -0.33333333333333331};

// ---------------------------------------------------------------------------
// Giles' inverse Poisson CDF (reference Numerics.hpp:49-272; ACM TOMS Alg.
// 955) in FP64. Decides the integer N that fixes the CIR mixture of the NCI
// and Exact schemes, so it runs in double on both path precisions.
//
// poissinv_w takes W = Phi^-1(U) computed as +-Phi^-1(min(U, 1 - U)) by
// giles_w (only read when lam > 4): the Greek variants of one sample share
// their Poisson uniform, so they share W -- the costliest part of the
// method -- and each runs only its own asymptotic and summation steps.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double giles_w(double U) {
  const double V = 1.0 - U;
  const double W = normcdfinv_d(fmin(U, V));
  return U > V ? -W : W;
}

// Giles' central asymptotic estimate (lam > 4, |W| < 3): continuous
// quantile S and its error bound Del. Shared by the full inversion and the
// lockstep variants' shortcut so both take identical arithmetic.
__device__ __forceinline__ void giles_asym_S(double lam, double ilam, double W, double &S, double &Del) {
  const double sl = sqrt(lam);
  S = sl * W + (1.0 / 3.0 + (1.0 / 6.0) * W * W) * (1.0 - W / (12.0 * sl));
  Del = 1.0 / 160.0;
  Del = 1.0 / 80.0 + Del * (W * W);
  Del = 1.0 / 40.0 + Del * (W * W);
  Del = Del * ilam;
  S = lam + (S + Del);
}

// What the FP32 first pass (poisson_first32) knows about its count N, for
// the bumped slots' counts (bumped_count). valid: N came from the bottom-up
// sum, with up = F(N) - U >= 0, lo = U - F(N-1) > 0 (lo = U for N = 0), the
// probabilities p(N), p(N-1).
// The pass returns such an N only where both margins exceed kSumTol, far
// above the FP32 sum's own error. asym: N = floor(S) from the central
// asymptotic branch, with S (within s_tol of the FP64 value) and its error
// term Del. Wf: that pass's W = Phi^-1(U) in FP32. Stored in FP32: the
// pass's error bounds are far above FP32 rounding of these values.
constexpr float kSumTol = 3e-5f;
struct PoisMargin {
  bool valid, asym;
  float up, lo, p_hi, p_lo;
  float S, Del, s_tol;
  float Wf;
};

// DEFER: instead of running the rare branches (the tail Newton iteration for
// |W| >= 3 and the Temme / Stirling correction near an integer boundary), set
// *rare and return early; the caller then runs the full inversion for those
// cases (poissinv), so the result is the same.
template <bool DEFER = false>
__device__ __forceinline__ double poissinv_w(double lam, double U, double W, bool *rare = nullptr) {
  if (U <= 0.0) return 0.0;
  if (U >= 1.0) return __builtin_nan("");
  const double V = 1.0 - U;
  const double ilam = 1.0 / lam;
  double X = 0.0, Del = 0.0;
  if (lam > 4.0) {
    double S;
    if (fabs(W) < 3.0) {
      giles_asym_S(lam, ilam, W, S, Del);
    } else {
      // uncontracted, as the oracle: (1 - R) + R log R vanishes as R -> 1
#pragma clang fp contract(off)
      if constexpr (DEFER) {
        *rare = true;
        return 0.0;
      }
      S = W / sqrt(lam);
      double R = 1.0 + S;
      if (R < 0.1) R = 0.1;
      double Rp;
      do {
        const double Tl = log(R);
        Rp = R;
        double S2 = sqrt(2.0 * ((1.0 - R) + R * Tl));
        if (R < 1.0) S2 = -S2;
        R = Rp - (S2 - S) * S2 / Tl;
        if (R < 0.1 * Rp) R = 0.1 * Rp;
      } while (fabs(R - Rp) > 1e-8);
      const double Tl = log(R);
      S = lam * R + log(sqrt(2.0 * R * ((1.0 - R) + R * Tl)) / fabs(R - 1.0)) / Tl;
      S = S - 0.0218 / (S + 0.065 * lam);
      Del = 0.01 / S;
      S = S + Del;
    }
    X = floor(S);
    if (S > 10.0 && S < X + 2.0 * Del) {
      if constexpr (DEFER) {
        *rare = true;
        return X;
      }
      if (X > 0.5 * lam && X < 2.0 * lam) {
        const double Xi = 1.0 / X;
        double Eta;
        {
          // no FMA contraction here: fma(-X, ilam, 1) keeps the unrounded
          // product, so at X = lam (integer intensities) 1 - Eta came out
          // -5.6e-17 instead of 0 and the square root NaN, which skipped
          // the correction below and returned N + 1
#pragma clang fp contract(off)
          Eta = X * ilam;
          Eta = sqrt(2.0 * (1.0 - Eta + Eta * log(Eta)) / Eta);
        }
        if (X > lam) Eta = -Eta;
        // Temme expansion coefficients (kTemmeB1 / kTemmeB0): odd / even
        // Horner chains in 1/X
        const double *b1 = kTemmeB1, *b0 = kTemmeB0;
        double B1 = b1[0], Sm = B1, B0 = b0[0];
        Sm = B0 + Sm * Eta;
        double mult = 25.0;
#pragma unroll
        for (int j = 1; j < 13; ++j) {
          B1 = b1[j] + mult * B1 * Xi;
          Sm = B1 + Sm * Eta;
          mult -= 1.0;
          B0 = b0[j] + mult * B0 * Xi;
          Sm = B0 + Sm * Eta;
          mult -= 1.0;
        }
        Sm = Sm / (1.0 + B1 * Xi);
        Sm = Sm * exp(-0.5 * X * Eta * Eta) / sqrt(2.0 * 3.141592653589793 * X);
        if (X < lam) {
          Sm += 0.5 * erfc(Eta * sqrt(0.5 * X));
          if (Sm > U) X -= 1.0;
        } else {
          Sm -= 0.5 * erfc(-Eta * sqrt(0.5 * X));
          if (Sm > -V) X -= 1.0;
        }
      } else {
        double Xi = 1.0 / X;
        double Sm = -(691.0 / 360360.0);
        Sm = (1.0 / 1188.0) + Sm * Xi * Xi;
        Sm = -(1.0 / 1680.0) + Sm * Xi * Xi;
        Sm = (1.0 / 1260.0) + Sm * Xi * Xi;
        Sm = -(1.0 / 360.0) + Sm * Xi * Xi;
        Sm = (1.0 / 12.0) + Sm * Xi * Xi;
        Sm = Sm * Xi;
        Sm = (X - lam) - X * log(X * ilam) - Sm;
        if (X < lam) {
          double Tt = exp(-0.5 * Sm);
          Sm = 1.0 - Tt * (U * Tt) * sqrt(2.0 * 3.141592653589793 * Xi) * lam;
          Tt = 1.0;
          Xi = X;
          for (int i = 1; i < 50; ++i) {
            Xi -= 1.0;
            Tt *= Xi * ilam;
            Sm += Tt;
          }
          if (Sm > 0.0) X -= 1.0;
        } else {
          const double Tt = exp(-0.5 * Sm);
          Sm = 1.0 - Tt * (V * Tt) * sqrt(2.0 * 3.141592653589793 * X);
          Xi = X;
          for (int i = 0; i < 50; ++i) {
            Xi += 1.0;
            Sm = Sm * Xi * ilam + 1.0;
          }
          if (Sm < 0.0) X -= 1.0;
        }
      }
    }
  }
  if (X < 10.0) {
    X = 0.0;
    double Tt = exp(0.5 * lam);
    Del = 0.0;
    if (U > 0.5) Del = Tt * (1e-13 * Tt);
    double Sm = 1.0 - Tt * (U * Tt) + Del;
    while (Sm < 0.0) {
      X += 1.0;
      Tt = X * ilam;
      Del = Tt * Del;
      Sm = Tt * Sm + 1.0;
    }
    if (Sm < 2.0 * Del) {
      Del = 1e13 * Del;
      Tt = 1e17 * Del;
      Del = V * Del;
      while (Del < Tt) {
        X += 1.0;
        Del *= X * ilam;
      }
      Sm = Del;
      Tt = 1.0;
      while (Sm > 0.0) {
        Tt *= X * ilam;
        Sm -= Tt;
        X -= 1.0;
      }
    }
  }
  return X;
}

__device__ static __noinline__ double poissinv(double lam, double U) {
  return poissinv_w(lam, U, (lam > 4.0 && U > 0.0 && U < 1.0) ? giles_w(U) : 0.0);
}

// ---------------------------------------------------------------------------
// Poisson counts, FP32 first pass. Giles' method returns the exact quantile
// of Poisson(lam) at U except where U lies within rounding (~1e-12) of a CDF
// value. This pass finds that quantile in FP32 with an error bound and
// returns it only where U clears the bound; elsewhere, and in the method's
// rare branches (|W| >= 3, the Temme / Stirling correction band), it
// returns -1 and the FP64 method decides. So its counts are the method's.
//  - lam > 4: the central asymptotic S in FP32 from W = Phi^-1(U) in FP32;
//    |S32 - S64| < tol = 1e-4 + 2e-6 lam (FP32 rounding of terms up to
//    lam + 3 sqrt(lam), and W's FP32 error times sqrt(lam)). floor(S) when
//    S >= 10 + tol and S clears the integers and the correction band
//    [X, X + 2 Del) by tol; the bottom-up sum when S < 10 - tol (the method
//    sums from 0 there too).
//  - lam <= 4, or S < 10: F(n) = sum of p(j), p(j) = p(j-1) lam / j,
//    p(0) = e^-lam, in FP32 (relative error below 1e-5 for the counts, at
//    most ~25, this branch reaches); n when F(n) - U and U - F(n-1) both
//    exceed 3e-5.
// ---------------------------------------------------------------------------
__constant__ static const float kInvN[48] = {0.0f, 1.0f / 1, 1.0f / 2, 1.0f / 3, 1.0f / 4, 1.0f / 5, 1.0f / 6, 1.0f / 7, 1.0f / 8, 1.0f / 9, 1.0f / 10, 1.0f / 11, 1.0f / 12, 1.0f / 13, 1.0f / 14, 1.0f / 15, 1.0f / 16, 1.0f / 17, 1.0f / 18, 1.0f / 19, 1.0f / 20, 1.0f / 21, 1.0f / 22, 1.0f / 23, 1.0f / 24, 1.0f / 25, 1.0f / 26, 1.0f / 27, 1.0f / 28, 1.0f / 29, 1.0f / 30, 1.0f / 31, 1.0f / 32, 1.0f / 33, 1.0f / 34, 1.0f / 35, 1.0f / 36, 1.0f / 37, 1.0f / 38, 1.0f / 39, 1.0f / 40, 1.0f / 41, 1.0f / 42, 1.0f / 43, 1.0f / 44, 1.0f / 45, 1.0f / 46, 1.0f / 47};

__device__ __forceinline__ float giles_w32(double U) {
  const double V = 1.0 - U;
  const float W = normcdfinv_f(fmin(U, V));
  return U > V ? -W : W;
}

__device__ __forceinline__ void giles_asym_S32(float lam, float W, float &S, float &Del) {
  const float sl = __builtin_amdgcn_sqrtf(lam);
  S = sl * W + (1.0f / 3.0f + (1.0f / 6.0f) * W * W) * (1.0f - W * __builtin_amdgcn_rcpf(12.0f * sl));
  Del = (((1.0f / 160.0f) * (W * W) + 1.0f / 80.0f) * (W * W) + 1.0f / 40.0f) * __builtin_amdgcn_rcpf(lam);
  S = lam + (S + Del);
}

// The asymptotic branch's count in FP32 (lam > 4, |W| < 3), or -1 when S is
// within tol of 10, an integer or the correction band.
__device__ __forceinline__ int giles_asym32(double lam, float W, float &S, float &Del, float &tol) {
  const float lf = static_cast<float>(lam);
  giles_asym_S32(lf, W, S, Del);
  tol = 1e-4f + 2e-6f * lf;
  const float X = floorf(S);
  const float fr = S - X;
  if (S >= 10.0f + tol && fr >= 2.0f * Del + tol && fr <= 1.0f - tol) return static_cast<int>(X);
  return -1;
}

// Giles' tail branch (lam > 4, |W| >= 3) in FP32: Newton's iteration for
// the R with eta(R) = W / sqrt(lam), eta(R) = sign(R - 1) sqrt(2 (1 - R +
// R log R)), then S and Del as the FP64 method forms them. Returns floor(S)
// when S clears 10, the integers and the correction band [X, X + 2 Del) by
// tol, else -1 with S set for the caller's S < 10 - tol test (the bottom-up
// sum). tol = 1e-3 + 1e-5 lam is 12x the largest |S32 - S64| found with W
// from the FP32 normal quantile and every FP32 logarithm, square root and
// reciprocal perturbed by up to 2 ulp, over lam in (4, 3200] and |W| in
// [3, 9] where S64 >= 9 (tests/test_poisson_fp32_pass.py). Two cases need no
// iteration, both S64 < 10: eta > -sqrt(2) for every R > 0, so below
// W / sqrt(lam) = -1.3 the method's iteration has no root (or one below
// R = 0.03) and its S stays below 2; and for W < 0 (R < 1) its S is below
// lam R + 0.4875 + 0.033 (the log-ratio term lies in [1/3, 0.4875] there),
// under 10 for lam <= 9.4. An iteration that has not converged to 1e-6
// (relative) after 12 steps is left to the FP64 method.
__device__ __forceinline__ int giles_tail32(double lam, float W, float &S, float &tol) {
  const float lf = static_cast<float>(lam);
  constexpr float kLn2 = 0.69314718055994531f;
  tol = 1e-3f + 1e-5f * lf;
  const float s = W * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(lf));
  S = 0.0f;
  if (s < -1.3f || (W < 0.0f && lf <= 9.4f)) return -1;
  float R = fmaxf(1.0f + s, 0.1f);
  bool conv = false;
  for (int it = 0; it < 12; ++it) {
    const float Tl = kLn2 * __builtin_amdgcn_logf(R);
    const float Rp = R;
    float S2 = __builtin_amdgcn_sqrtf(2.0f * ((1.0f - R) + R * Tl));
    if (R < 1.0f) S2 = -S2;
    R = Rp - (S2 - s) * S2 * __builtin_amdgcn_rcpf(Tl);
    if (R < 0.1f * Rp) R = 0.1f * Rp;
    if (fabsf(R - Rp) <= 1e-6f * R) {
      conv = true;
      break;
    }
  }
  if (!conv) {
    S = 10.0f;
    return -1;
  }
  const float Tl = kLn2 * __builtin_amdgcn_logf(R);
  const float a = __builtin_amdgcn_sqrtf(2.0f * R * ((1.0f - R) + R * Tl));
  S = lf * R + kLn2 * __builtin_amdgcn_logf(a * __builtin_amdgcn_rcpf(fabsf(R - 1.0f))) * __builtin_amdgcn_rcpf(Tl);
  S = S - 0.0218f * __builtin_amdgcn_rcpf(S + 0.065f * lf);
  const float Del = 0.01f * __builtin_amdgcn_rcpf(S);
  S = S + Del;
  const float X = floorf(S);
  const float fr = S - X;
  if (S >= 10.0f + tol && fr >= 2.0f * Del + tol && fr <= 1.0f - tol) return static_cast<int>(X);
  return -1;
}

// Giles' correction near an integer boundary (S > 10 in [X, X + 2 Del),
// 0.5 lam < X < 2 lam) in FP32: Temme's expansion Sm of the CDF at X - 1 as
// the FP64 method forms it (the kTemmeB1 / kTemmeB0 chains), compared with
// U (X < lam) or -V (X >= lam) to choose X - 1 or X. The FP32 form takes
// r - 1 = (X - lam) / lam from FP64 and 1 - r + r log r from its series in
// r - 1 below |r - 1| = 1/4, so eta carries no cancellation. Returns the
// count where Sm clears U (or -V) by 4e-6, 6x the largest |Sm32 - Sm64|
// found with every FP32 transcendental perturbed by up to 2 ulp (erfc by
// 4) over X >= 10, lam in (5, 3200] (tests/test_poisson_fp32_pass.py); else
// -1.
__device__ __forceinline__ int giles_temme32(double lam, double U, int X) {
  constexpr float kLn2 = 0.69314718055994531f;
  const double Xd = static_cast<double>(X);
  const float d = static_cast<float>((Xd - lam) * (1.0 / lam));
  const float r = 1.0f + d;
  float g;
  if (fabsf(d) < 0.25f) {
    // sum over k >= 2 of (-1)^k d^k / (k (k - 1)), to k = 14
    float p = 1.0f / 182.0f;
    p = p * d - 1.0f / 156.0f;
    p = p * d + 1.0f / 132.0f;
    p = p * d - 1.0f / 110.0f;
    p = p * d + 1.0f / 90.0f;
    p = p * d - 1.0f / 72.0f;
    p = p * d + 1.0f / 56.0f;
    p = p * d - 1.0f / 42.0f;
    p = p * d + 1.0f / 30.0f;
    p = p * d - 1.0f / 20.0f;
    p = p * d + 1.0f / 12.0f;
    p = p * d - 1.0f / 6.0f;
    p = p * d + 1.0f / 2.0f;
    g = d * d * p;
  } else {
    g = r * (kLn2 * __builtin_amdgcn_logf(r)) - d;
  }
  float eta = __builtin_amdgcn_sqrtf(2.0f * g * __builtin_amdgcn_rcpf(r));
  if (Xd > lam) eta = -eta;
  const float Xf = static_cast<float>(X);
  const float xi = __builtin_amdgcn_rcpf(Xf);
  float B1 = static_cast<float>(kTemmeB1[0]), Sm = B1, B0 = static_cast<float>(kTemmeB0[0]);
  Sm = B0 + Sm * eta;
  float mult = 25.0f;
#pragma unroll
  for (int j = 1; j < 13; ++j) {
    B1 = static_cast<float>(kTemmeB1[j]) + mult * B1 * xi;
    Sm = B1 + Sm * eta;
    mult -= 1.0f;
    B0 = static_cast<float>(kTemmeB0[j]) + mult * B0 * xi;
    Sm = B0 + Sm * eta;
    mult -= 1.0f;
  }
  Sm = Sm * __builtin_amdgcn_rcpf(1.0f + B1 * xi);
  Sm = Sm * expf(-0.5f * Xf * eta * eta) *
       __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(6.28318530717958648f * Xf));
  const float z = eta * __builtin_amdgcn_sqrtf(0.5f * Xf);
  constexpr double kTol = 4e-6;
  if (Xd < lam) {
    const double c = static_cast<double>(Sm + 0.5f * erfcf(z));
    if (c > U + kTol) return X - 1;
    if (c < U - kTol) return X;
  } else {
    const double c = static_cast<double>(Sm - 0.5f * erfcf(-z));
    const double mV = U - 1.0;
    if (c > mV + kTol) return X - 1;
    if (c < mV - kTol) return X;
  }
  return -1;
}

// The bottom-up sum's count in FP32, or -1 when U lies within 3e-5 of F(n)
// or F(n-1) (or no count below 48 reaches U). The loop steps all active
// lanes together, so j is wave-uniform and 1/j a scalar load.
__device__ __forceinline__ int pois_sum32(double lam, double U, PoisMargin *mg) {
  const float lf = static_cast<float>(lam), uf = static_cast<float>(U);
  float p = expf(-lf), F = p, Fp = 0.0f, pp = 0.0f;
  int n = 0;
  bool run = F < uf;
  for (int j = 1; j < 48; ++j) {
    if (!__any(run)) break;
    if (run) {
      pp = p;
      Fp = F;
      p = p * lf * kInvN[j];
      F += p;
      n = j;
      run = F < uf;
    }
  }
  const float up = F - uf, lo = uf - Fp;
  if (run || !(up > kSumTol) || !(n == 0 || lo > kSumTol)) return -1;
  if (mg) {
    mg->valid = true;
    mg->up = up;
    mg->lo = n > 0 ? lo : uf;
    mg->p_hi = p;
    mg->p_lo = n > 0 ? pp : 0.0f;
  }
  return n;
}

// The count Giles' method returns for (lam, U), 0 <= U < 1, from the FP32
// first pass, or -1 (undecided). W = giles_w32(U), read when lam > 4. mg
// receives the margins bumped_count reads. TAIL: also the tail branch
// (giles_tail32) and the Temme correction (giles_temme32); without them
// those cases are left undecided, so that the lockstep slots of such a
// sample go to the wave-packed pass together.
template <bool TAIL = false>
__device__ __forceinline__ int poisson_first32(double lam, double U, float W, PoisMargin *mg) {
  if (mg) {
    mg->valid = false;
    mg->asym = false;
    mg->Wf = W;
  }
  if (U <= 0.0) return 0;
  if (lam > 4.0) {
    float S, Del, tol;
    if constexpr (TAIL) {
      // the method's branch is decided on FP64 W (|W| < 3): W32 is within
      // 1e-5 of it
      if (fabsf(W) > 3.001f) {
        const int N = giles_tail32(lam, W, S, tol);
        if (N >= 0) return N;
        if (!(S < 10.0f - tol)) return -1;
        return pois_sum32(lam, U, mg);
      }
    }
    if (!(fabsf(W) < 2.999f)) return -1;
    const int N = giles_asym32(lam, W, S, Del, tol);
    if (N >= 0) {
      if (mg) {
        mg->asym = true;
        mg->S = S;
        mg->Del = Del;
        mg->s_tol = tol;
      }
      return N;
    }
    if constexpr (TAIL) {
      // S64 > 10 inside the correction band, with X = floor(S64) and the
      // band's edges cleared by tol: the Temme branch, which the FP64 method
      // takes for 0.5 lam < X < 2 lam. At X = 10 a step down to 9 sends the
      // method to its bottom-up sum, whose count pois_sum32 finds.
      const float X = floorf(S), fr = S - X;
      if (S >= 10.0f + tol && fr >= tol && fr <= 2.0f * Del - tol) {
        const double Xd = static_cast<double>(X);
        if (!(Xd > 0.5 * lam && Xd < 2.0 * lam)) return -1;
        const int n = giles_temme32(lam, U, static_cast<int>(X));
        if (n == 9) return pois_sum32(lam, U, nullptr);
        return n;
      }
    }
    if (!(S < 10.0f - tol)) return -1;
  }
  return pois_sum32(lam, U, mg);
}

// Poisson(mean) by sequential inversion on one uniform (Boost's method for
// means below 10; JumpPolicies.hpp:60,116).
__device__ __forceinline__ int poisson_count(double mean, double exp_mean, double u) {
  double p = exp_mean;
  int x = 0;
  while (u > p) {
    u -= p;
    ++x;
    p = mean * p / x;
    if (p == 0.0) break;
  }
  return x;
}

// Jump count of one Poisson(mean) draw from one uniform. Below mean 10 the
// sequential inversion above, as Boost does (JumpPolicies.hpp:48,60); from
// 10 on Boost switches to transformed rejection (PTRD, a variable number of
// draws), replaced here by the exact inverse CDF (Giles' method, the NCI
// scheme's Poisson quantile) of the same single uniform: a different draw
// sequence, the same distribution.
__device__ static __noinline__ double poissinv(double lam, double U);
__device__ __forceinline__ int jump_count(double mean, double exp_mean, double u) {
  if (mean < 10.0) return poisson_count(mean, exp_mean, u);
  return static_cast<int>(poissinv(mean, u));
}

// Exponential(rate) by inversion (replaces Boost's ziggurat: one draw each).
template <typename R, typename Gen>
__device__ __forceinline__ R exponential(R rate, Gen &g) {
  return static_cast<R>(-log1p(-uniform(g))) / rate;
}

// ---------------------------------------------------------------------------
// Regularised incomplete gamma and its inverse (chi-squared quantile of the
// NCI scheme: table rows and the out-of-table fallback, VolSchemes.hpp:282,
// 368-369). Power series below a+1, Legendre continued fraction (Lentz)
// above; Halley steps inside a shrinking bracket.
// ---------------------------------------------------------------------------
// lga1 = lgamma(a + 1)
__device__ static __noinline__ double igam_series(double a, double x, double lga1) {
  double sum = 1.0, term = 1.0, ap = a;
  for (int n = 1; n < 1000; ++n) {
    ap += 1.0;
    term *= x / ap;
    sum += term;
    if (term <= sum * 1e-16) break;
  }
  return exp(a * log(x) - x - lga1) * sum;
}

// lga = lgamma(a)
__device__ static __noinline__ double igamc_cf(double a, double x, double lga) {
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) <= 1e-15) break;  // a few ulps of 1, above FMA noise
  }
  return exp(a * log(x) - x - lga) * h;
}

// x with P(a, x) = p, given lga = lgamma(a) and q = 1 - p. The upper
// branch (p > 1/2) is solved on Q(a, x) = q with q passed separately, so a
// tail probability far below 2^-53 keeps its precision. The device lgamma
// costs thousands of cycles, so callers pass it (the path kernels read it
// from a per-table list of lgamma(d/2 + N)). Accurate to ~1e-14 for
// p >= 2^-64, the smallest nonzero uniform the engine draws; far below that
// (p < ~1e-60 at a > 1) the starting point falls outside the bracket logic's
// reach and the result is not accurate.
__device__ static __noinline__ double gamma_inv(double a, double p, double q, double lga) {
  if (p <= 0.0) return 0.0;
  if (q <= 0.0) return __builtin_inf();
  const double lga1 = lga + log(a);  // lgamma(a + 1)
  const bool upper = p > 0.5;
  double x;
  if (a > 1.0) {
    const double z = upper ? -normcdfinv_d(q) : normcdfinv_d(p);
    const double t = 1.0 / (9.0 * a);
    const double w = 1.0 - t + z * sqrt(t);
    x = (w > 0.0) ? a * w * w * w : 0.5 * a;
  } else {
    const double xl = exp((log(p) + lga1) / a);
    if (!upper && xl < 1.0) {
      x = xl;
    } else {
      double xx = -log(q) - lga;
      if (xx < 1.0) xx = 1.0;
      for (int i = 0; i < 4; ++i) xx = -log(q) - lga + (a - 1.0) * log(xx);
      x = (xx > 0.0) ? xx : 1.0;
    }
  }
  if (!(x > 0.0)) x = 1e-300;
  double lo = 0.0, hi = __builtin_inf();
  for (int it = 0; it < 64; ++it) {
    double f;
    if (upper) f = q - ((x < a + 1.0) ? 1.0 - igam_series(a, x, lga1) : igamc_cf(a, x, lga));
    else f = ((x < a + 1.0) ? igam_series(a, x, lga1) : 1.0 - igamc_cf(a, x, lga)) - p;
    if (f > 0.0) hi = x; else lo = x;
    if (f == 0.0) break;
    const double pdf = exp((a - 1.0) * log(x) - x - lga);
    double xn = -1.0;
    if (pdf > 0.0 && isfinite(pdf)) {
      const double r = f / pdf;
      double den = 1.0 - 0.5 * r * ((a - 1.0) / x - 1.0);
      if (den <= 0.1) den = 1.0;
      xn = x - r / den;
      // converged: a step this small can round onto the bracket end x
      // itself, which must not be taken for a step out of the bracket
      if (fabs(xn - x) <= 1e-14 * xn) {
        if (xn > lo && xn < hi) x = xn;
        break;
      }
    }
    if (!(xn > lo && xn < hi)) {
      if (isinf(hi)) xn = (x > 0.0 ? 2.0 * x : 1.0);
      else if (lo > 0.0 && hi / lo > 4.0) xn = sqrt(lo * hi);
      else xn = 0.5 * (lo + hi);
    }
    const double dx = fabs(xn - x);
    x = xn;
    // Halley converges cubically: a step below 1e-14 x leaves an error far
    // below one ulp; the bracket test ends oscillation at the rounding floor
    if (dx <= 1e-14 * x || (hi - lo) <= 1e-14 * x) break;
  }
  return x;
}

// chi-squared quantile at probability p, lga = lgamma(dof / 2)
__device__ static double chi2_quantile_lg(double dof, double p, double lga) {
  return 2.0 * gamma_inv(0.5 * dof, p, 1.0 - p, lga);
}

// chi-squared quantile at upper-tail probability qt (P(X > x) = qt)
__device__ static double chi2_quantile_upper_lg(double dof, double qt, double lga) {
  return 2.0 * gamma_inv(0.5 * dof, 1.0 - qt, qt, lga);
}

__device__ static double chi2_quantile(double dof, double p) { return chi2_quantile_lg(dof, p, lgamma(0.5 * dof)); }

}  // namespace myql
