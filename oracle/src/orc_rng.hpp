// ORACLE (test infrastructure only): CPU FP64 restatement of the random
// number layer under the reference's Monte Carlo path. Only tests/, the
// smoke check and bench.py's cpu_baseline leg may load it.
//
// Reference call sites this follows (the reference uses Boost.Random, which
// is not vendored in the reference tree; Boost 1.88.0 per Dockerfile:19-20):
//   MonteCarloPricer.hpp:86,191       boost::random::xoshiro256pp(seed + tid*999983)
//   ASVJstepper.hpp:137               uniform_real_distribution<double>(0,1)
//   JumpPolicies.hpp:38-39,48,86-89   poisson_distribution<int,double>, uniform_01,
//                                     exponential_distribution
// Published algorithms restated here:
//   xoshiro256++ (Blackman & Vigna 2018, public domain reference code):
//     out = rotl(s0 + s3, 23) + s0, xorshift state update with rotl(s3, 45).
//   splitmix64 seeding of the 256-bit state (Boost's xoshiro seed(u64)).
//   Jump-ahead by 2^128: characteristic polynomial P(x) of the state
//   transition found by Berlekamp-Massey over GF(2); J_k = x^(2^(128+k)) mod P.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

// Op counting (build with -DORC_COUNT_OPS: lib/libmyql_oracle_ops.so). Each
// executed block adds its floating-point operation count -- an FMA counts 2;
// an add, mul, div, sqrt, exp, log or erfc counts 1 -- to g_ops. The counts
// are for the reference's algorithm as written (every Greek variant replays
// its own path); integer RNG work is not counted. Single-threaded use only.
// Set-up work (the NCI quantile tables) runs under OpsPause and is not
// counted.
#ifdef ORC_COUNT_OPS
inline uint64_t g_ops = 0;
inline bool g_ops_on = true;
#define ORC_OPS(n) (::orc::g_ops += ::orc::g_ops_on ? static_cast<uint64_t>(n) : 0u)
struct OpsPause {
  bool was = g_ops_on;
  OpsPause() { g_ops_on = false; }
  ~OpsPause() { g_ops_on = was; }
};
#else
#define ORC_OPS(n) ((void)0)
struct OpsPause {};
#endif

struct Xoshiro {
  uint64_t s[4];
};

static inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

static inline uint64_t xnext(Xoshiro &g) {
  const uint64_t out = rotl64(g.s[0] + g.s[3], 23) + g.s[0];
  const uint64_t t = g.s[1] << 17;
  g.s[2] ^= g.s[0];
  g.s[3] ^= g.s[1];
  g.s[1] ^= g.s[2];
  g.s[0] ^= g.s[3];
  g.s[2] ^= t;
  g.s[3] = rotl64(g.s[3], 45);
  return out;
}

// splitmix64: state += golden gamma, then the two xor-shift-multiply rounds.
static inline uint64_t splitmix_next(uint64_t &x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline Xoshiro seed_state(uint64_t seed) {
  Xoshiro g;
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) g.s[i] = splitmix_next(x);
  return g;
}

// uniform_real_distribution<double>(0,1) / uniform_01<double> over a 64-bit
// integer engine: u = double(x) / 2^64 (round to nearest), redrawn while
// u == 1.0 (x >= 2^64 - 2^10).
#ifndef ORC_NO_DRAW_LOG
// Debug exports only (oracle_debug_*_path): when set, every uniform this
// thread draws is appended, so a test can replay the draws of a path. The
// CPU-baseline and op-counting builds compile it out (-DORC_NO_DRAW_LOG).
inline thread_local std::vector<double> *g_draw_log = nullptr;
#endif

static inline double uniform(Xoshiro &g) {
  for (;;) {
    const double u = static_cast<double>(xnext(g)) * 0x1p-64;
    ORC_OPS(1);
    if (u < 1.0) {
#ifndef ORC_NO_DRAW_LOG
      if (g_draw_log) g_draw_log->push_back(u);
#endif
      return u;
    }
  }
}

static inline double draw_index() {
#ifndef ORC_NO_DRAW_LOG
  return g_draw_log ? static_cast<double>(g_draw_log->size()) : -1.0;
#else
  return -1.0;
#endif
}

// ---------------------------------------------------------------------------
// GF(2) polynomials of degree < 512, bit i = coefficient of x^i.
// ---------------------------------------------------------------------------
struct Poly {
  uint64_t w[8];
};

static inline int poly_bit(const Poly &p, int i) { return (p.w[i >> 6] >> (i & 63)) & 1; }
static inline void poly_flip(Poly &p, int i) { p.w[i >> 6] ^= (1ull << (i & 63)); }

// Characteristic polynomial of the xoshiro256 state transition, from the
// sequence of bit 0 of s[0] (Berlekamp-Massey over 512 terms).
static inline Poly char_poly() {
  const int N = 512;
  std::vector<int> seq(N);
  Xoshiro g = seed_state(0x1234567ull);
  for (int n = 0; n < N; ++n) {
    seq[n] = static_cast<int>(g.s[0] & 1);
    xnext(g);
  }
  std::vector<int> C(N + 1, 0), B(N + 1, 0), Tm(N + 1, 0);
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < N; ++n) {
    int d = seq[n];
    for (int i = 1; i <= L; ++i) d ^= C[i] & seq[n - i];
    if (!d) {
      ++m;
    } else if (2 * L <= n) {
      Tm = C;
      for (int i = 0; i + m <= N; ++i) C[i + m] ^= B[i];
      L = n + 1 - L;
      B = Tm;
      m = 1;
    } else {
      for (int i = 0; i + m <= N; ++i) C[i + m] ^= B[i];
      ++m;
    }
  }
  // P(x) = x^L C(1/x): coefficient of x^j is C[L - j].
  Poly P;
  std::memset(&P, 0, sizeof(P));
  for (int j = 0; j <= L; ++j)
    if (C[L - j]) poly_flip(P, j);
  return P;
}

static inline int poly_degree(const Poly &p) {
  for (int i = 511; i >= 0; --i)
    if (poly_bit(p, i)) return i;
  return -1;
}

// r^2 mod P, with deg r < deg P = 256.
static inline Poly poly_sqr_mod(const Poly &r, const Poly &P) {
  Poly s;
  std::memset(&s, 0, sizeof(s));
  for (int i = 0; i < 256; ++i)
    if (poly_bit(r, i)) poly_flip(s, 2 * i);
  const int dP = poly_degree(P);
  for (int d = 2 * 255; d >= dP; --d) {
    if (!poly_bit(s, d)) continue;
    for (int j = 0; j <= dP; ++j)
      if (poly_bit(P, j)) poly_flip(s, d - dP + j);
  }
  return s;
}

// x * r mod P
static inline Poly poly_mulx_mod(const Poly &r, const Poly &P) {
  Poly s;
  std::memset(&s, 0, sizeof(s));
  for (int i = 0; i < 511; ++i)
    if (poly_bit(r, i)) poly_flip(s, i + 1);
  const int dP = poly_degree(P);
  if (poly_bit(s, dP))
    for (int j = 0; j <= dP; ++j)
      if (poly_bit(P, j)) poly_flip(s, j);
  return s;
}

// x^(2^e) mod P
static inline Poly x_pow2_mod(int e, const Poly &P) {
  Poly r;
  std::memset(&r, 0, sizeof(r));
  poly_flip(r, 1);
  for (int i = 0; i < e; ++i) r = poly_sqr_mod(r, P);
  return r;
}

// state <- J(T) state, J given by its 256 low coefficients (xoshiro's jump
// routine: accumulate the states whose coefficient bit is set).
static inline void apply_poly(Xoshiro &g, const uint64_t J[4]) {
  uint64_t a[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int b = 0; b < 64; ++b) {
      if (J[i] & (1ull << b))
        for (int k = 0; k < 4; ++k) a[k] ^= g.s[k];
      xnext(g);
    }
  for (int k = 0; k < 4; ++k) g.s[k] = a[k];
}

// Jump polynomials J_k = x^(2^(128+k)) mod P for k = 0..nk-1.
static inline std::vector<std::array<uint64_t, 4>> jump_polys(int nk);

}  // namespace orc

#include <array>
namespace orc {
static inline std::vector<std::array<uint64_t, 4>> jump_polys(int nk) {
  const Poly P = char_poly();
  std::vector<std::array<uint64_t, 4>> out(nk);
  Poly r = x_pow2_mod(128, P);
  for (int k = 0; k < nk; ++k) {
    for (int i = 0; i < 4; ++i) out[k][i] = r.w[i];
    r = poly_sqr_mod(r, P);
  }
  return out;
}
}  // namespace orc
