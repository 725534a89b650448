// ORACLE (test infrastructure only): CPU FP64 restatement of the special
// functions on the reference's Monte Carlo path.
//
//   normcdfinv   reference math/Numerics.hpp:12-47 (Acklam's rational
//                approximation of the inverse normal CDF, coefficients as
//                published by P. J. Acklam, 2003).
//   poissinvcdf  reference math/Numerics.hpp:49-272 (M. Giles, "Algorithm
//                955: approximation of the inverse Poisson cumulative
//                distribution function", ACM TOMS 2016), with Acklam in place
//                of the normal quantile exactly as the reference does.
//   poisson_inversion  Boost poisson_distribution<int,double> for means below
//                10 (sequential inversion): JumpPolicies.hpp:48,60.
//   jump_count   the jump counts: poisson_inversion below mean 10; from 10
//                on Boost's PTRD rejection sampler is replaced by
//                poissinvcdf of one uniform (documented deviation, pinned
//                against scipy's Poisson quantile and pmf).
//   exponential  Boost exponential_distribution is a ziggurat whose tables
//                are not available here; replaced by inversion
//                -log1p(-u)/rate, one uniform per variate (documented
//                deviation, pinned statistically).
//   chi2_quantile  boost::math::quantile(chi_squared(d), p)
//                (VolSchemes.hpp:282,368-369) restated as 2 * P^-1(d/2, p),
//                P the regularised lower incomplete gamma function, inverted
//                by Halley's method on a bracketed interval.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

#include "orc_rng.hpp"

namespace orc {

static inline double normcdfinv(double p) {
  static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                              -2.759285104469687e+02, 1.383577518672690e+02,
                              -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                              -1.556989798598866e+02, 6.680131188771972e+01,
                              -1.328068155288572e+01};
  static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                              -2.400758277161838e+00, -2.549732539343734e+00,
                              4.374664141464968e+00,  2.938163982698783e+00};
  static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                              2.445134137142996e+00, 3.754408661907416e+00};
  const double lo = 0.02425, hi = 1.0 - 0.02425;
  if (p > lo && p < hi) {
    const double q = p - 0.5, r = q * q;
    const double num = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q;
    const double den = ((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0;
    ORC_OPS(24);  // q, r, 5+5 FMA, *q, divide
    return num / den;
  }
  const double q = (p <= lo) ? p : 1.0 - p;
  const double r = std::sqrt(-2.0 * std::log(q));
  const double x = (((((c[0] * r + c[1]) * r + c[2]) * r + c[3]) * r + c[4]) * r + c[5]) /
                   ((((d[0] * r + d[1]) * r + d[2]) * r + d[3]) * r + 1.0);
  ORC_OPS(23);  // 1-p, log, *-2, sqrt, 5+4 FMA, divide
  return (p <= lo) ? x : -x;
}

// Giles' inverse Poisson CDF. Returns the smallest integer n with
// F(n) >= U (as a double); 0 for U <= 0, NaN for U >= 1.
static inline double poissinvcdf(double lam, double U) {
  if (U <= 0.0) return 0.0;
  if (U >= 1.0) return std::numeric_limits<double>::quiet_NaN();
  const double V = 1.0 - U;
  const double ilam = 1.0 / lam;
  double X = 0.0, S = 0.0, Del = 0.0;
  ORC_OPS(2);

  if (lam > 4.0) {
    double W = normcdfinv(U < V ? U : V);
    if (U > V) W = -W;
    if (std::fabs(W) < 3.0) {
      // central region: Giles' asymptotic expansion in W
      const double sl = std::sqrt(lam);
      S = sl * W + (1.0 / 3.0 + (1.0 / 6.0) * W * W) * (1.0 - W / (12.0 * sl));
      Del = 1.0 / 160.0;
      Del = 1.0 / 80.0 + Del * (W * W);
      Del = 1.0 / 40.0 + Del * (W * W);
      Del = Del * ilam;
      S = lam + (S + Del);
      ORC_OPS(19);
    } else {
      // tails: Newton iteration for r in sqrt(2(1 - r + r log r)) = S
      S = W / std::sqrt(lam);
      double R = 1.0 + S;
      if (R < 0.1) R = 0.1;
      double Rprev;
      do {
        const double Tl = std::log(R);
        Rprev = R;
        double S2 = std::sqrt(2.0 * ((1.0 - R) + R * Tl));
        if (R < 1.0) S2 = -S2;
        R = Rprev - (S2 - S) * S2 / Tl;
        if (R < 0.1 * Rprev) R = 0.1 * Rprev;
        ORC_OPS(12);
      } while (std::fabs(R - Rprev) > 1e-8);
      const double Tl = std::log(R);
      S = lam * R + std::log(std::sqrt(2.0 * R * ((1.0 - R) + R * Tl)) /
                             std::fabs(R - 1.0)) / Tl;
      S = S - 0.0218 / (S + 0.065 * lam);
      Del = 0.01 / S;
      S = S + Del;
      ORC_OPS(20);
    }
    X = std::floor(S);
    ORC_OPS(1);

    if (S > 10.0 && S < X + 2.0 * Del) {
      if (X > 0.5 * lam && X < 2.0 * lam) {
        // Temme's uniform asymptotic expansion of the Poisson CDF
        const double Xi = 1.0 / X;
        double Eta = X * ilam;
        Eta = std::sqrt(2.0 * (1.0 - Eta + Eta * std::log(Eta)) / Eta);
        if (X > lam) Eta = -Eta;
        ORC_OPS(10);
        // Coefficient table of the expansion (Giles, Alg. 955), applied as
        // the two interleaved Horner recurrences B0/B1 in 1/X.
        static const double cb1[13] = {
            8.0995211567045583e-16, -5.1391118342426808e-16, -1.3923887224181616e-13,
            1.1004392031956284e-13, 2.4361948020667402e-11,  -2.5514193994946487e-11,
            -4.3820360184533521e-09, 6.7078535434015332e-09, 8.2967113409530833e-07,
            -2.1854485106799979e-06, -0.00017875514403292177, 0.0011574074074074078,
            0.083333333333333329};
        static const double cb0[13] = {
            -1.9752288294349411e-15, 2.8534893807047458e-14, 3.3717632624009806e-13,
            -5.0276692801141763e-12, -5.8307721325504166e-11, 9.1476995822367933e-10,
            1.0261809784240299e-08,  -1.7665952736826086e-07, -1.8540622107151585e-06,
            3.9192631785224383e-05,  0.00035273368606701921,  -0.014814814814814815,
            -0.33333333333333331};
        double B1 = cb1[0];
        double Sm = B1;
        double B0 = cb0[0];
        Sm = B0 + Sm * Eta;
        // multipliers 25, 24, ..., 2 alternate between B1 and B0
        int mult = 25;
        for (int j = 1; j < 13; ++j) {
          B1 = cb1[j] + mult * B1 * Xi;
          Sm = B1 + Sm * Eta;
          --mult;
          B0 = cb0[j] + mult * B0 * Xi;
          Sm = B0 + Sm * Eta;
          --mult;
          ORC_OPS(10);
        }
        Sm = Sm / (1.0 + B1 * Xi);
        Sm = Sm * std::exp(-0.5 * X * Eta * Eta) / std::sqrt(2.0 * 3.141592653589793 * X);
        ORC_OPS(19);  // Horner start, normalisation, exp/sqrt factor, erfc term
        if (X < lam) {
          Sm += 0.5 * std::erfc(Eta * std::sqrt(0.5 * X));
          if (Sm > U) X -= 1.0;
        } else {
          Sm -= 0.5 * std::erfc(-Eta * std::sqrt(0.5 * X));
          if (Sm > -V) X -= 1.0;
        }
      } else {
        // Stirling series for log(X!) and a 50-term sum down or up
        double Xi = 1.0 / X;
        double Sm = -(691.0 / 360360.0);
        Sm = (1.0 / 1188.0) + Sm * Xi * Xi;
        Sm = -(1.0 / 1680.0) + Sm * Xi * Xi;
        Sm = (1.0 / 1260.0) + Sm * Xi * Xi;
        Sm = -(1.0 / 360.0) + Sm * Xi * Xi;
        Sm = (1.0 / 12.0) + Sm * Xi * Xi;
        Sm = Sm * Xi;
        Sm = (X - lam) - X * std::log(X * ilam) - Sm;
        ORC_OPS(24);
        if (X < lam) {
          double Tt = std::exp(-0.5 * Sm);
          Sm = 1.0 - Tt * (U * Tt) * std::sqrt(2.0 * 3.141592653589793 * Xi) * lam;
          Tt = 1.0;
          Xi = X;
          for (int i = 1; i < 50; ++i) {
            Xi -= 1.0;
            Tt *= Xi * ilam;
            Sm += Tt;
          }
          ORC_OPS(6 + 49 * 4);
          if (Sm > 0.0) X -= 1.0;
        } else {
          const double Tt = std::exp(-0.5 * Sm);
          Sm = 1.0 - Tt * (V * Tt) * std::sqrt(2.0 * 3.141592653589793 * X);
          Xi = X;
          for (int i = 0; i < 50; ++i) {
            Xi += 1.0;
            Sm = Sm * Xi * ilam + 1.0;
          }
          ORC_OPS(6 + 50 * 4);
          if (Sm < 0.0) X -= 1.0;
        }
      }
    }
  }

  if (X < 10.0) {
    // bottom-up summation of the scaled CDF
    X = 0.0;
    double Tt = std::exp(0.5 * lam);
    Del = 0.0;
    if (U > 0.5) Del = Tt * (1e-13 * Tt);
    double Sm = 1.0 - Tt * (U * Tt) + Del;
    ORC_OPS(7);
    while (Sm < 0.0) {
      X += 1.0;
      Tt = X * ilam;
      Del = Tt * Del;
      Sm = Tt * Sm + 1.0;
      ORC_OPS(5);
    }
    if (Sm < 2.0 * Del) {
      // top-down summation when rounding leaves the answer ambiguous
      Del = 1e13 * Del;
      Tt = 1e17 * Del;
      Del = V * Del;
      while (Del < Tt) {
        X += 1.0;
        Del *= X * ilam;
        ORC_OPS(3);
      }
      Sm = Del;
      Tt = 1.0;
      while (Sm > 0.0) {
        Tt *= X * ilam;
        Sm -= Tt;
        X -= 1.0;
        ORC_OPS(4);
      }
    }
  }
  return X;
}

// Poisson(mean) by sequential inversion: one uniform; p_0 = e^-mean,
// subtract p_x while u exceeds it, p_{x+1} = mean * p_x / (x+1).
static inline int poisson_inversion(double mean, double exp_mean, Xoshiro &g) {
  double p = exp_mean;
  int x = 0;
  double u = uniform(g);
  while (u > p) {
    u = u - p;
    ++x;
    p = mean * p / x;
    ORC_OPS(3);
    if (p == 0.0) break;  // u left above a rounded-down total: stop
  }
  return x;
}

// Jump count: sequential inversion below mean 10 (Boost's branch there);
// from 10 on Boost's PTRD rejection sampler is replaced by the exact inverse
// CDF of the same single uniform (poissinvcdf): the same distribution, a
// different draw sequence (the engine does the same).
static inline int jump_count(double mean, double exp_mean, Xoshiro &g) {
  if (mean < 10.0) return poisson_inversion(mean, exp_mean, g);
  ORC_OPS(60);
  return static_cast<int>(poissinvcdf(mean, uniform(g)));
}

static inline double exponential(double rate, Xoshiro &g) {
  ORC_OPS(3);
  return -std::log1p(-uniform(g)) / rate;
}

// ---------------------------------------------------------------------------
// Regularised incomplete gamma P(a,x), Q(a,x) and the inverse of P.
// ---------------------------------------------------------------------------
static inline double lgam(double a) { return std::lgamma(a); }

// P by its power series: x^a e^-x / Gamma(a+1) * sum x^n / ((a+1)...(a+n))
static inline double gamma_p_series(double a, double x) {
  double sum = 1.0, term = 1.0, ap = a;
  for (int n = 1; n < 1000; ++n) {
    ap += 1.0;
    term *= x / ap;
    sum += term;
    ORC_OPS(4);
    if (term <= sum * 1e-16) break;
  }
  ORC_OPS(7);
  return std::exp(a * std::log(x) - x - lgam(a + 1.0)) * sum;
}

// Q by the Legendre continued fraction, evaluated with Lentz's method.
static inline double gamma_q_cf(double a, double x) {
  const double tiny = 1e-300;
  double b = x + 1.0 - a;
  double c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    ORC_OPS(12);
    if (std::fabs(del - 1.0) <= 1e-15) break;  // a few ulps of 1, above FMA noise
  }
  ORC_OPS(7);
  return std::exp(a * std::log(x) - x - lgam(a)) * h;
}

static inline double gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  return (x < a + 1.0) ? gamma_p_series(a, x) : 1.0 - gamma_q_cf(a, x);
}
static inline double gamma_q(double a, double x) {
  if (x <= 0.0) return 1.0;
  return (x < a + 1.0) ? 1.0 - gamma_p_series(a, x) : gamma_q_cf(a, x);
}

// x with P(a, x) = p, q = 1 - p passed separately. The lower tail is solved
// on P, the upper tail (p > 0.5) on Q = q so that a tail probability far
// below 2^-53 keeps its precision. Pinned against scipy for p >= 2^-64 (the
// smallest nonzero uniform, tests/golden chi2_low_tail); far below that
// (p < ~1e-60 at a > 1) it is not accurate.
static inline double gamma_inv(double a, double p, double q) {
  if (p <= 0.0) return 0.0;
  if (q <= 0.0) return std::numeric_limits<double>::infinity();
  const bool upper = p > 0.5;
  const double lg_a = lgam(a);
  double x;
  if (a > 1.0) {
    const double z = upper ? -normcdfinv(q) : normcdfinv(p);
    const double t = 1.0 / (9.0 * a);
    const double w = 1.0 - t + z * std::sqrt(t);
    x = (w > 0.0) ? a * w * w * w : 0.5 * a;
  } else {
    // small shape: P ~ x^a / Gamma(a+1) near 0, Q ~ x^(a-1) e^-x / Gamma(a)
    const double x_lo = std::exp((std::log(p) + lgam(a + 1.0)) / a);
    if (!upper && x_lo < 1.0) {
      x = x_lo;
    } else {
      double xx = -std::log(q) - lg_a;
      if (xx < 1.0) xx = 1.0;
      for (int i = 0; i < 4; ++i) xx = -std::log(q) - lg_a + (a - 1.0) * std::log(xx);
      x = (xx > 0.0) ? xx : 1.0;
    }
  }
  if (!(x > 0.0)) x = 1e-300;
  double lo = 0.0, hi = std::numeric_limits<double>::infinity();
  for (int it = 0; it < 64; ++it) {
    // residual f (increasing in x) and pdf
    double f;
    if (upper) f = q - gamma_q(a, x);  // = P - p computed through Q
    else f = gamma_p(a, x) - p;
    if (f > 0.0) hi = x; else lo = x;
    if (f == 0.0) break;
    const double lpdf = (a - 1.0) * std::log(x) - x - lg_a;
    const double pdf = std::exp(lpdf);
    double xn;
    if (pdf > 0.0 && std::isfinite(pdf)) {
      const double r = f / pdf;
      const double curv = (a - 1.0) / x - 1.0;
      double denom = 1.0 - 0.5 * r * curv;
      if (denom <= 0.1) denom = 1.0;  // fall back to Newton
      xn = x - r / denom;
      // converged: a step this small can round onto the bracket end x
      // itself, which must not be taken for a step out of the bracket
      if (std::fabs(xn - x) <= 1e-14 * xn) {
        if (xn > lo && xn < hi) x = xn;
        ORC_OPS(16);
        break;
      }
    } else {
      xn = -1.0;
    }
    if (!(xn > lo && xn < hi)) {
      // outside the bracket: bisect (geometrically when it spans decades)
      if (std::isinf(hi)) xn = (x > 0.0 ? 2.0 * x : 1.0);
      else if (lo > 0.0 && hi / lo > 4.0) xn = std::sqrt(lo * hi);
      else xn = 0.5 * (lo + hi);
    }
    ORC_OPS(16);  // residual, log-pdf, exp, Halley step
    const double dx = std::fabs(xn - x);
    x = xn;
    // Halley converges cubically: a step below 1e-14 x leaves an error far
    // below one ulp; the bracket test ends oscillation at the rounding floor
    if (dx <= 1e-14 * x || (hi - lo) <= 1e-14 * x) break;
  }
  return x;
}

static inline double gamma_p_inv(double a, double p) { return gamma_inv(a, p, 1.0 - p); }

// chi-squared quantile at upper-tail probability qt
static inline double chi2_quantile_upper(double dof, double qt) { return 2.0 * gamma_inv(0.5 * dof, 1.0 - qt, qt); }

static inline double chi2_quantile(double dof, double p) {
  return 2.0 * gamma_p_inv(0.5 * dof, p);
}

}  // namespace orc
