// Gil-Pelaez Fourier pricer of European options under the ASVJ models, on
// the GPU: the reference's FourierPricer (pricers/fourier/FourierPricer.hpp
// :34-407) with its GilPelaezKernel integrands (engines/fourier/kernels/
// GilPelaezKernel.hpp:32-119) and characteristic exponents (models/asvj/core/
// AffineTraits.hpp:18-79, policies/CFPolicies.hpp:9-60).
//
// One wavefront prices one option (strike of one job). The integrals over
// u in [0, U] are taken by composite 16-point Gauss-Legendre on 64 panels of
// geometrically growing width (ratio 1.08, as the reference's chunk grid),
// one panel per lane; U is the first of start_bound * 1.5^k (k < 20; the
// reference's find_upper_bound, FourierEngine.hpp:100-115) where both price
// integrands fall below 1e-12, times 1.5. Every integrand of a node -- price,
// Dx, Dxx, vega per factor, theta, the cross kernels -- shares that node's
// two characteristic-exponent evaluations (u - i for P1, u for P2), which
// the reference recomputes per integral; the lane partials are summed
// across the wave. All arithmetic is FP64.
#include <hip/hip_runtime.h>

#include "fourier_launch.hpp"

namespace myql {
namespace {

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cx(double r, double i = 0.0) { return {r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx operator*(double s, cplx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ cplx operator/(cplx a, cplx b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, d = b.re + b.im * t;
    return {(a.re + a.im * t) / d, (a.im - a.re * t) / d};
  }
  const double t = b.re / b.im, d = b.re * t + b.im;
  return {(a.re * t + a.im) / d, (a.im * t - a.re) / d};
}
__device__ __forceinline__ cplx cexp(cplx a) {
  double s, c;
  sincos(a.im, &s, &c);
  const double e = exp(a.re);
  return {e * c, e * s};
}
__device__ __forceinline__ cplx clog(cplx a) { return {log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }
// principal square root
__device__ __forceinline__ cplx csqrt(cplx a) {
  const double m = hypot(a.re, a.im);
  if (m == 0.0) return {0.0, 0.0};
  if (a.re >= 0.0) {
    const double t = sqrt(0.5 * (m + a.re));
    return {t, a.im / (2.0 * t)};
  }
  const double t = sqrt(0.5 * (m - a.re));
  return {fabs(a.im) / (2.0 * t), copysign(t, a.im)};
}

// Heston factor (Albrecher's form, AffineTraits.hpp:18-36) and its v0 and T
// derivatives (:41-80), at complex u.
struct HestonCF {
  cplx psi, dv0, dT;
};
__device__ __forceinline__ HestonCF heston_cf(const FourierModel &m, int f, cplx u, double t, bool derivs) {
  if (t < 1e-8) return {cx(0), cx(0), cx(0)};
  const double kappa = m.h[f][0], theta = m.h[f][1], sigma = m.h[f][2], rho = m.h[f][3], v0 = m.h[f][4];
  const double s2 = sigma * sigma;
  const cplx iu = {-u.im, u.re};
  const cplx dp = (rho * sigma) * iu;
  const cplx km = dp - cx(kappa);
  const cplx D = csqrt(km * km + s2 * (iu + u * u));
  const cplx A = cx(kappa) - dp - D;
  const cplx g = A / (cx(kappa) - dp + D);
  const cplx eT = cexp((-t) * D);
  const cplx den = cx(1.0) - g * eT;
  HestonCF r;
  r.psi = (kappa * theta / s2) * (t * A - 2.0 * clog(den / (cx(1.0) - g))) + (v0 / s2) * (A * (cx(1.0) - eT) / den);
  if (derivs) {
    r.dv0 = A * (cx(1.0) - eT) / (s2 * den);
    r.dT = (kappa * theta / s2) * (A - 2.0 * (g * D * eT / den)) + (v0 / s2) * (A * D * eT * (cx(1.0) - g) / (den * den));
  } else {
    r.dv0 = r.dT = cx(0);
  }
  return r;
}

// compensated jump exponent per unit time (CFPolicies.hpp:9-60)
__device__ __forceinline__ cplx jump_exponent(const FourierModel &m, cplx u) {
  const cplx iu = {-u.im, u.re};
  if (m.jump_kind == 1) {
    const double lam = m.jump[0], mu = m.jump[1], delta = m.jump[2];
    const double k = exp(mu + 0.5 * delta * delta) - 1.0;
    return lam * (cexp(mu * iu - (0.5 * delta * delta) * (u * u)) - cx(1.0) - k * iu);
  }
  if (m.jump_kind == 2) {
    const double lam = m.jump[0], p = m.jump[1], e1 = m.jump[2], e2 = m.jump[3];
    const double k = p * e1 / (e1 - 1.0) + (1.0 - p) * e2 / (e2 + 1.0) - 1.0;
    return lam * (cx(p * e1) / (cx(e1) - iu) + cx((1.0 - p) * e2) / (cx(e2) + iu) - cx(1.0) - k * iu);
  }
  return cx(0);
}

// log-characteristic exponent of the martingale part and its derivatives
struct CF {
  cplx psi, dv0[2], dT;
};
__device__ __forceinline__ CF model_cf(const FourierModel &m, cplx u, double t, bool derivs) {
  CF c;
  const cplx jmp = jump_exponent(m, u);
  if (m.nf == 0) {
    const cplx iu = {-u.im, u.re};
    const cplx q = iu + u * u;
    c.psi = (-0.5 * m.vol * m.vol * t) * q + t * jmp;
    c.dv0[0] = (-m.vol * t) * q;
    c.dv0[1] = cx(0);
    c.dT = (-0.5 * m.vol * m.vol) * q + jmp;
    return c;
  }
  c.psi = t * jmp;
  c.dT = jmp;
  c.dv0[1] = cx(0);
  for (int f = 0; f < m.nf; ++f) {
    const HestonCF h = heston_cf(m, f, u, t, derivs);
    c.psi = c.psi + h.psi;
    c.dv0[f] = h.dv0;
    c.dT = c.dT + h.dT;
  }
  return c;
}

// integrals per option (each already divided by pi): for j = 1 (P1, shift
// u - i) and j = 2 (P2, shift u): price, Dx, Dxx, vega[2], theta, vega-Dx[2],
// theta-Dx; rho integrals are T x Dx
enum { I_P, I_DX, I_DXX, I_V0, I_V1, I_T, I_VDX0, I_VDX1, I_TDX, N_INT };

__device__ __forceinline__ double wave_sum(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__constant__ double kGLx[8] = {0.09501250983763745, 0.2816035507792589, 0.45801677765722737, 0.6178762444026438,
                               0.755404408355003,   0.8656312023878318, 0.9445750230732326,  0.9894009349916499};
__constant__ double kGLw[8] = {0.18945061045506859, 0.1826034150449236,  0.16915651939500262, 0.14959598881657676,
                               0.12462897125553403, 0.09515851168249259, 0.062253523938647706, 0.027152459411754037};

// integrand magnitude of the two price kernels at real u (bound search)
__device__ double price_magnitude(const FourierOption &o, double u) {
  const FourierModel &m = o.model;
  double mx = 0.0;
  for (int j = 0; j < 2; ++j) {
    const cplx us = j == 0 ? cx(u, -1.0) : cx(u, 0.0);
    const CF c = model_cf(m, us, o.T, false);
    const cplx e = cexp(c.psi + cx(0.0, u * (o.drift - o.logK)));
    const cplx v = e / cx(0.0, u);
    mx = fmax(mx, fabs(v.re));
  }
  return mx;
}

__global__ void __launch_bounds__(256) fourier_kernel(const FourierOption *__restrict__ opts, int n_opts,
                                                      double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int o_idx = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (o_idx >= n_opts) return;  // whole waves exit together
  const FourierOption &o = opts[o_idx];
  const bool derivs = o.mode == 2;

  // upper bound: lanes 0..19 test start * 1.5^k, the first passing k wins
  double v = o.start_bound * pow(1.5, static_cast<double>(lane));
  const bool ok = lane < 20 && price_magnitude(o, v) < 1e-12;
  const unsigned long long b = __ballot(ok);
  const int k = b ? __ffsll(static_cast<long long>(b)) - 1 : 19;
  const double U = 1.5 * o.start_bound * pow(1.5, static_cast<double>(k));

  // panel [a, a + w) of lane: widths w_p = w_0 * 1.08^p, sum = U
  const double ratio = 1.08;
  const double w0 = U * (ratio - 1.0) / (pow(ratio, 64.0) - 1.0);
  const double a = w0 * (pow(ratio, static_cast<double>(lane)) - 1.0) / (ratio - 1.0);
  const double w = w0 * pow(ratio, static_cast<double>(lane));
  const double mid = a + 0.5 * w, half = 0.5 * w;

  double acc[2][N_INT];
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < N_INT; ++i) acc[j][i] = 0.0;
  const double rq = o.T > 1e-12 ? o.drift / o.T : 0.0;
  for (int n = 0; n < 16; ++n) {
    const double x = n < 8 ? -kGLx[7 - n] : kGLx[n - 8];
    const double wt = (n < 8 ? kGLw[7 - n] : kGLw[n - 8]) * half;
    const double u = mid + half * x;
    const cplx iu = {0.0, u};
    for (int j = 0; j < 2; ++j) {
      const cplx us = j == 0 ? cx(u, -1.0) : cx(u, 0.0);
      const CF c = model_cf(o.model, us, o.T, derivs);
      const cplx base = cexp(c.psi + cx(0.0, u * (o.drift - o.logK)));
      const cplx bi = base / iu;
      acc[j][I_P] += wt * bi.re;
      acc[j][I_DX] += wt * base.re;
      acc[j][I_DXX] += wt * (base * iu).re;
      if (derivs) {
        const cplx tf = c.dT + rq * iu;
        acc[j][I_V0] += wt * (bi * c.dv0[0]).re;
        acc[j][I_V1] += wt * (bi * c.dv0[1]).re;
        acc[j][I_T] += wt * (bi * tf).re;
        acc[j][I_VDX0] += wt * (base * c.dv0[0]).re;
        acc[j][I_VDX1] += wt * (base * c.dv0[1]).re;
        acc[j][I_TDX] += wt * (base * tf).re;
      }
    }
  }
  const double inv_pi = 0.31830988618379067154;
  double *dst = out + static_cast<size_t>(o_idx) * 2 * N_INT;
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < N_INT; ++i) {
      const double s = wave_sum(acc[j][i]);
      if (lane == 0) dst[j * N_INT + i] = s * inv_pi;
    }
}

}  // namespace

int fourier_n_integrals() { return 2 * N_INT; }

hipError_t launch_fourier(const FourierOption *opts, int n_opts, double *out, hipStream_t stream) {
  if (n_opts <= 0) return hipSuccess;
  const int waves_per_block = 4;
  const unsigned nb = static_cast<unsigned>((n_opts + waves_per_block - 1) / waves_per_block);
  hipLaunchKernelGGL(fourier_kernel, dim3(nb), dim3(64 * waves_per_block), 0, stream, opts, n_opts, out);
  return hipGetLastError();
}

}  // namespace myql
