// Host side of the MI355X Monte Carlo engine: the C-ABI of
// include/myql_capi.h, device context and caches, and the translation of a
// MonteCarloPricer::calculate call (reference MonteCarloPricer.hpp:95-185)
// into kernel parameter blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <bitset>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "myql_capi.h"
#include "dev_math.hpp"
#include "engine_types.hpp"
#include "mc_launch.hpp"
#include "mlmc_kernel.hpp"
#include "cv_launch.hpp"
#include "cv_types.hpp"
#include "fourier_launch.hpp"

using namespace myql;

namespace {

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw Fail(e_ == hipErrorOutOfMemory ? MYQL_E_OOM : MYQL_E_HIP,                      \
                 std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

// growable device buffer
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  void reserve(size_t bytes) {
    if (bytes <= n) return;
    // a queued kernel may still read the old buffer: wait for the device
    // before freeing it (growth only, so this is rare)
    if (p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    HIPCHECK(hipMalloc(&p, bytes));
    n = bytes;
  }
  ~DevBuf() {
    if (p) {
      (void)hipDeviceSynchronize();  // an error path may leave work queued
      (void)hipFree(p);
    }
  }
  template <typename T>
  T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------------------
// Jump-ahead polynomials of xoshiro256: J_k(x) = x^(2^(128+k)) mod P(x),
// P the characteristic polynomial of the state transition, found from a bit
// sequence of the generator by the Berlekamp-Massey recurrence over GF(2).
// ---------------------------------------------------------------------------
using Bits = std::bitset<512>;

void host_step(uint64_t s[4]) {
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 45) | (s[3] >> 19);
}

Bits transition_charpoly() {
  const int n = 512;
  uint64_t s[4] = {0x243F6A8885A308D3ull, 0x13198A2E03707344ull, 0xA4093822299F31D0ull,
                   0x082EFA98EC4E6C89ull};
  Bits seq;
  for (int i = 0; i < n; ++i) {
    seq[i] = (s[3] >> 7) & 1;
    host_step(s);
  }
  Bits c, b;  // connection polynomial and its last copy
  c[0] = b[0] = 1;
  int L = 0, shift = 1;
  for (int i = 0; i < n; ++i) {
    bool disc = seq[i];
    for (int j = 1; j <= L; ++j) disc ^= (c[j] && seq[i - j]);
    if (!disc) {
      ++shift;
      continue;
    }
    const Bits prev = c;
    c ^= (b << shift);
    if (2 * L <= i) {
      L = i + 1 - L;
      b = prev;
      shift = 1;
    } else {
      ++shift;
    }
  }
  Bits P;  // reciprocal of c: x^L c(1/x)
  for (int j = 0; j <= L; ++j) P[j] = c[L - j];
  return P;
}

Bits sqr_mod(const Bits &a, const Bits &P, int deg) {
  Bits r;
  for (int i = 0; i < deg; ++i)
    if (a[i]) r[2 * i] = 1;
  for (int i = 2 * deg - 2; i >= deg; --i)
    if (r[i]) r ^= (P << (i - deg));
  return r;
}

std::vector<uint64_t> host_jump_polys(int nk) {
  const Bits P = transition_charpoly();
  int deg = 511;
  while (deg > 0 && !P[deg]) --deg;
  Bits r;
  r[1] = 1;
  for (int i = 0; i < 128; ++i) r = sqr_mod(r, P, deg);
  std::vector<uint64_t> out(4 * nk);
  for (int k = 0; k < nk; ++k) {
    for (int w = 0; w < 4; ++w) {
      uint64_t x = 0;
      for (int b = 0; b < 64; ++b)
        if (r[64 * w + b]) x |= (1ull << b);
      out[4 * k + w] = x;
    }
    r = sqr_mod(r, P, deg);
  }
  return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
namespace myql {

// n 32-bit words set to zero (queue counters)
__global__ void zero_words_kernel(unsigned *__restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__device__ void apply_jump(Rng &g, const uint64_t *__restrict__ J) {
  Rng a = {0, 0, 0, 0};
  for (int w = 0; w < 4; ++w) {
    const uint64_t word = J[w];
    for (int b = 0; b < 64; ++b) {
      if ((word >> b) & 1ull) {
        a.s0 ^= g.s0;
        a.s1 ^= g.s1;
        a.s2 ^= g.s2;
        a.s3 ^= g.s3;
      }
      step_state(g);
    }
  }
  g = a;
}

__global__ void stream_states_kernel(uint64_t seed, int layout, uint64_t S, const uint64_t *__restrict__ polys,
                                     uint64_t *__restrict__ out) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  Rng g;
  if (layout == MYQL_STREAMS_SEEDED) {
    g = seed_rng(seed + s * 999983ull);
  } else {
    g = seed_rng(seed);
    for (int k = 0; k < 64; ++k) {
      if ((s >> k) == 0) break;
      if ((s >> k) & 1ull) apply_jump(g, polys + 4 * k);
    }
  }
  out[4 * s + 0] = g.s0;
  out[4 * s + 1] = g.s1;
  out[4 * s + 2] = g.s2;
  out[4 * s + 3] = g.s3;
}

// Row k holds chi2 quantiles for dof d + 2k at u = i*dx (VolSchemes.hpp:267-290);
// lg[k] = lgamma(d/2 + k).
__global__ void nci_quantiles_kernel(double d, int rows, int grid, double dx, const double *__restrict__ lg,
                                     double *__restrict__ y) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<int64_t>(rows) * grid) return;
  const int k = static_cast<int>(e / grid), i = static_cast<int>(e % grid);
  y[e] = chi2_quantile_lg(d + 2.0 * k, static_cast<double>(i) * dx, lg[k]);
}

// Fritsch-Butland slopes (VolSchemes.hpp:292-310) for rows k <= n_fb, exact
// slopes dq/du = 1/pdf(q) above, in the interleaved {y, dx*m} layout the path
// kernels gather from.
template <typename R>
__global__ void nci_pack_kernel(const double *__restrict__ y, int rows, int grid, double dx, double inv_dx,
                                double d, int n_fb, const double *__restrict__ lg, R *__restrict__ tab) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<int64_t>(rows) * grid) return;
  const int i = static_cast<int>(e % grid);
  const int k = static_cast<int>(e / grid);
  const double *r = y + (e - i);
  double m;
  if (k > n_fb && i > 0) {
    const double dof = d + 2.0 * k;
    m = exp(-((0.5 * dof - 1.0) * log(r[i]) - 0.5 * r[i] - 0.5 * dof * 0.69314718055994531 - lg[k]));
  } else if (i == 0) m = (r[1] - r[0]) * inv_dx;
  else if (i == grid - 1) m = (r[i] - r[i - 1]) * inv_dx;
  else {
    const double d1 = (r[i] - r[i - 1]) * inv_dx, d2 = (r[i + 1] - r[i]) * inv_dx;
    m = (d1 * d2 > 0.0) ? 2.0 * d1 * d2 / (d1 + d2) : 0.0;
  }
  tab[2 * e] = static_cast<R>(r[i]);
  tab[2 * e + 1] = static_cast<R>(dx * m);
}

// Sum of the unit partials of each job in unit order.
__global__ void reduce_partials_kernel(const double *__restrict__ partial, double *__restrict__ sums,
                                       const long long *__restrict__ base, const int *__restrict__ n_slots,
                                       const int *__restrict__ nk, const int *__restrict__ out_off,
                                       int n_jobs) {
  const int job = blockIdx.y;
  if (job >= n_jobs) return;
  const int width = 2 * NSTAT * nk[job];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= width) return;
  double t = 0.0;
  const double *p = partial + base[job] + e;
  for (int i = 0; i < n_slots[job]; ++i) t += p[static_cast<size_t>(i) * width];
  sums[out_off[job] + e] = t;
}


// Upper-tail rows: quantiles at tail probability e^-w, w = w0 + j hw, and
// hw * dq/dw = hw e^-w / pdf(q).
template <typename R>
__global__ void nci_tail_kernel(double d, int rows, int pts, double w0, double hw, const double *__restrict__ lg,
                                R *__restrict__ tail) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<int64_t>(rows) * pts) return;
  const int k = static_cast<int>(e / pts), j = static_cast<int>(e % pts);
  const double dof = d + 2.0 * k;
  const double w = w0 + j * hw;
  const double qt = exp(-w);
  const double x = chi2_quantile_upper_lg(dof, qt, lg[k]);
  const double lpdf = (0.5 * dof - 1.0) * log(x) - 0.5 * x - 0.5 * dof * 0.69314718055994531 - lg[k];
  tail[2 * e] = static_cast<R>(x);
  tail[2 * e + 1] = static_cast<R>(hw * qt * exp(-lpdf));
}

__global__ void debug_draws_kernel(const uint64_t *__restrict__ states, uint64_t S, uint64_t n,
                                   uint64_t *__restrict__ out) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  Rng g = {states[4 * s], states[4 * s + 1], states[4 * s + 2], states[4 * s + 3]};
  for (uint64_t i = 0; i < n; ++i) out[s * n + i] = next_u64(g);
}

__global__ void debug_normcdfinv_kernel(const double *__restrict__ p, int64_t n, double *__restrict__ o64,
                                        float *__restrict__ o32) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  o64[i] = normcdfinv_d(p[i]);
  o32[i] = normcdfinv_f(p[i]);
}

__global__ void debug_poissinv_kernel(const double *__restrict__ lam, const double *__restrict__ u, int64_t n,
                                      double *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = poissinv(lam[i], u[i]);
}

// The path kernels' Poisson counts (FP32 first pass, bumped slots from slot
// 0's margins, the FP64 method on wave-packed leftovers) for NV slots that
// share one uniform; lanes past n run on placeholder inputs, since every
// lane of a wave takes part in the packing.
template <int NV>
__global__ void __launch_bounds__(BLOCK) debug_poisson_counts_kernel(const double *__restrict__ lam,
                                                                     const double *__restrict__ u, int64_t n,
                                                                     int32_t *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
  const bool live = i < n;
  double l[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) l[k] = live ? lam[i * NV + k] : 1.0;
  const double uu = live ? u[i] : 0.5;
  int Ns[NV];
  if constexpr (NV == 1) Ns[0] = poisson_count_packed(l[0], uu);
  else poisson_counts_lockstep<NV>([&](int k) { return l[k]; }, uu, Ns);
  if (live)
#pragma unroll
    for (int k = 0; k < NV; ++k) out[i * NV + k] = Ns[k];
}

// The quasi-Monte Carlo coordinates the path kernels draw: lane l of wave w
// (S / 64 waves) at iteration it < n_it reads point 64 (w n_it + it) + l
// through the same hi-table walk (qmc_hi_init, qmc_hi_step) and Qrng
// reads; out[i * dims + d] = coordinate d of point i < n_points.
__global__ void __launch_bounds__(BLOCK) debug_qmc_kernel(const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ lo,
                                                          const uint32_t *__restrict__ shift, int dims,
                                                          uint64_t n_waves, uint64_t n_it, uint64_t n_points,
                                                          uint32_t *__restrict__ out) {
  __shared__ uint32_t s_qhi[(BLOCK / 64) * QMC_DMAX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * (BLOCK / 64) + wave;
  if (w >= n_waves) return;  // whole waves exit together
  uint32_t *hi = s_qhi + wave * QMC_DMAX;
  const uint64_t h0 = w * n_it;
  qmc_hi_init(hi, V, shift, dims, h0, lane);
  wave_sync();
  for (uint64_t it = 0; it < n_it; ++it) {
    if (it > 0) {
      qmc_hi_step(hi, V, dims, h0 + it, lane);
      wave_sync();
    }
    const uint64_t i = 64 * (h0 + it) + static_cast<uint64_t>(lane);
    if (i < n_points) {
      Qrng q{hi, lo + lane, 0};
      for (int d = 0; d < dims; ++d) out[i * static_cast<uint64_t>(dims) + d] = qmc_next(q);
    }
    wave_sync();
  }
}

__global__ void debug_chi2_kernel(const double *__restrict__ dof, const double *__restrict__ p, int64_t n,
                                  double *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = chi2_quantile(dof[i], p[i]);
}

}  // namespace myql

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
namespace {
struct CvRecord;
}

struct myql_ctx {
  int device = 0;
  int n_cu = 256;
  size_t lds_max = 64 * 1024;  // dynamic LDS available to one block
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::vector<uint64_t> jump_polys;
  DevBuf d_polys;
  std::map<std::tuple<uint64_t, int, uint64_t>, std::unique_ptr<DevBuf>> streams;
  std::map<std::tuple<uint64_t, int, int, int>, std::unique_ptr<DevBuf>> nci;  // (d bits, rows, prec, n_fb)
  DevBuf jobs, units, partial, sums, meta, counter, tmp, accg;
  DevBuf mlmc_streams, mlmc_job;
  // launch shapes chosen by the autotuner per (precision, factor count,
  // lockstep width, path dependence), with the trial times behind them
  struct Tuned {
    myql::McShape shape;
    int gridf;
    std::string trials;
  };
  std::map<std::tuple<int, int, int, bool>, Tuned> tuned;
  DevBuf tjobs, tunits, tcounter;
  DevBuf vaccg;  // accumulators of fused (view) jobs
  DevBuf fopts, fout;  // Fourier pricer options and integrals
  // myql_mc_calculate_noalloc: this device's copy of the file-scope scratch
  // array (its device address, from hipGetSymbolAddress at myql_create), and
  // while such a call runs, the bump offset into it
  void *scratch = nullptr;
  bool noalloc = false;
  size_t scratch_off = 0;
  // learned control variate: recordings (their buffers grow and are reused
  // by later calls) and launch arrays
  std::vector<std::unique_ptr<CvRecord>> cv_recs;
  DevBuf cv_jobs, cv_sets, cv_tasks, cv_part, cv_sums, cv_coef;
  // terminal-jump tables E[P_k | S_T-], keyed by everything they depend on
  // (jump law and parameters, maturity, S0, payoff, type, smoothing, strikes):
  // a pricing call reuses the table its training call built
  std::map<std::vector<double>, std::unique_ptr<DevBuf>> cv_rb;
  // randomized QMC: Sobol' direction integers [32][QMC_DMAX] and the
  // lower-6-bit table [QMC_DMAX][64] (built at the first such call), and
  // the digital shifts of a call's replicates
  DevBuf qmc_V, qmc_lo, qmc_shift;
};

// Device memory of myql_mc_calculate_noalloc: stream states, job blocks,
// work units, queue counters, per-unit accumulators and partial sums of one
// call are carved from this array instead of being allocated. Each GPU the
// library runs on holds its own copy until the process ends; every such call
// on one GPU shares it, so calls that overlap on one GPU race.
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

// Called at the start of every pricing entry point: the stream-state and
// NCI-table caches are bounded between calls only, since the jobs of one
// call keep pointers into them until it returns.
void set_device(myql_ctx *c) {
  HIPCHECK(hipSetDevice(c->device));
  if (c->streams.size() > 16) c->streams.clear();
  if (c->nci.size() > 32) c->nci.clear();
  if (c->cv_rb.size() > 64) c->cv_rb.clear();
}

// Device memory of `bytes` for buffer b: b grown as needed, or, inside
// myql_mc_calculate_noalloc, the next 256-byte aligned piece of the scratch
// array (never an allocation there; too large a call is refused).
void *dev_mem(myql_ctx *c, DevBuf &b, size_t bytes) {
  if (!c->noalloc) {
    b.reserve(bytes);
    return b.p;
  }
  const size_t off = (c->scratch_off + 255) / 256 * 256;
  if (bytes > NOALLOC_BYTES || off > NOALLOC_BYTES - bytes)
    throw Fail(MYQL_E_INVALID_ARG, "myql_mc_calculate_noalloc: the call needs more device memory than the " +
                                       std::to_string(NOALLOC_BYTES >> 20) +
                                       " MiB scratch (fewer streams or strikes, or use myql_mc_calculate)");
  c->scratch_off = off + bytes;
  return static_cast<char *>(c->scratch) + off;
}

// S stream states for (seed, layout), cached per context (computed into the
// scratch, uncached, inside myql_mc_calculate_noalloc).
const uint64_t *stream_states(myql_ctx *c, uint64_t seed, int layout, uint64_t S) {
  if (c->noalloc) {
    DevBuf none;
    uint64_t *p = static_cast<uint64_t *>(dev_mem(c, none, S * 32));
    const int tb = 256;
    const unsigned nb = static_cast<unsigned>((S + tb - 1) / tb);
    hipLaunchKernelGGL(stream_states_kernel, dim3(nb), dim3(tb), 0, c->stream, seed, layout, S,
                       c->d_polys.as<uint64_t>(), p);
    HIPCHECK(hipGetLastError());
    return p;
  }
  auto key = std::make_tuple(seed, layout, S);
  auto it = c->streams.find(key);
  if (it != c->streams.end()) return it->second->as<uint64_t>();
  auto buf = std::make_unique<DevBuf>();
  buf->reserve(S * 32);
  const int tb = 256;
  const unsigned nb = static_cast<unsigned>((S + tb - 1) / tb);
  hipLaunchKernelGGL(stream_states_kernel, dim3(nb), dim3(tb), 0, c->stream, seed, layout, S,
                     c->d_polys.as<uint64_t>(), buf->as<uint64_t>());
  HIPCHECK(hipGetLastError());
  const uint64_t *p = buf->as<uint64_t>();
  c->streams[key] = std::move(buf);
  return p;
}

struct NciInfo {
  const void *tab;
  const void *tail;    // upper-tail rows
  const double *lg;    // lgamma(d/2 + N), N < NCI_LG
  const double *mean;  // per row, the mean of its quantile map (default semantics)
  int rows, grid;
  double dx, inv_dx;
};

constexpr int NCI_GRID = 10000;
constexpr int NCI_LG = 2048;
constexpr int NCI_TAIL = 512;
constexpr double NCI_TAIL_W1 = 38.0;
const double NCI_TAIL_W0 = std::log(1000.0);

// NCI rows for one workspace. Reference rule (VolSchemes.hpp:245-255):
// intensity at v_max = max(v0, 3 theta), n_max = ceil(l + 4 sqrt(l)). The
// default sizes the table at v = max(v0, 25 theta) with a 6 sqrt(l) margin so
// that the exact chi2 inversion above the last row is practically never run;
// both are clamped to [23, 1023].
int nci_nmax(const myql_heston_params &p, double dt, bool reference_rows) {
  const double v_max = std::max(p.v0, (reference_rows ? 3.0 : 25.0) * p.theta);
  const double ex = std::exp(-p.kappa * dt);
  const double psi = (std::fabs(p.kappa) < 1e-10) ? dt : (1.0 - ex) / p.kappa;
  const double lam_max = 2.0 * v_max * ex / (p.sigma * p.sigma * psi);
  double nm = std::ceil(lam_max + (reference_rows ? 4.0 : 6.0) * std::sqrt(lam_max));
  size_t n = static_cast<size_t>(nm);
  n = std::max<size_t>(n, 23);
  n = std::min<size_t>(n, 1023);
  return static_cast<int>(n);
}

// One buffer per (d, rows, precision, n_fb): main rows [rows x grid x 2],
// tail rows [rows x NCI_TAIL x 2] (R precision), then NCI_LG doubles of
// lgamma(d/2 + N), then the rows' means (the learned control variate's
// zero-mean variance increments).
NciInfo nci_table(myql_ctx *c, double d, int nmax, int prec, int n_fb) {
  const int rows = nmax + 1;
  uint64_t dbits;
  std::memcpy(&dbits, &d, 8);
  const double dx = 0.999 / static_cast<double>(NCI_GRID - 1);
  const double inv_dx = 1.0 / dx;
  const size_t esz = prec == MYQL_FP32 ? 4 : 8;
  const size_t main_b = static_cast<size_t>(rows) * NCI_GRID * 2 * esz;
  const size_t tail_b = static_cast<size_t>(rows) * NCI_TAIL * 2 * esz;
  auto info = [&](const DevBuf &b) {
    const char *p = static_cast<const char *>(b.p);
    return NciInfo{p, p + main_b, reinterpret_cast<const double *>(p + main_b + tail_b),
                   reinterpret_cast<const double *>(p + main_b + tail_b + NCI_LG * sizeof(double)), rows, NCI_GRID,
                   dx, inv_dx};
  };
  const auto key = std::make_tuple(dbits, rows, prec, n_fb);
  auto it = c->nci.find(key);
  if (it != c->nci.end()) return info(*it->second);
  const int64_t n = static_cast<int64_t>(rows) * NCI_GRID;
  c->tmp.reserve(n * sizeof(double));
  auto buf = std::make_unique<DevBuf>();
  buf->reserve(main_b + tail_b + (NCI_LG + rows) * sizeof(double));
  NciInfo ni = info(*buf);
  // lgamma(d/2 + N) for the quantile's incomplete-gamma prefactors
  std::vector<double> lgh(NCI_LG);
  for (int k = 0; k < NCI_LG; ++k) lgh[k] = std::lgamma(0.5 * d + k);
  double *d_lg = const_cast<double *>(ni.lg);
  HIPCHECK(hipMemcpyAsync(d_lg, lgh.data(), NCI_LG * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const int tb = 256;
  const unsigned nb = static_cast<unsigned>((n + tb - 1) / tb);
  hipLaunchKernelGGL(nci_quantiles_kernel, dim3(nb), dim3(tb), 0, c->stream, d, rows, NCI_GRID, dx, d_lg,
                     c->tmp.as<double>());
  HIPCHECK(hipGetLastError());
  const int64_t nt = static_cast<int64_t>(rows) * NCI_TAIL;
  const unsigned nbt = static_cast<unsigned>((nt + tb - 1) / tb);
  const double hw = (NCI_TAIL_W1 - NCI_TAIL_W0) / (NCI_TAIL - 1);
  if (prec == MYQL_FP32) {
    hipLaunchKernelGGL(nci_pack_kernel<float>, dim3(nb), dim3(tb), 0, c->stream, c->tmp.as<double>(), rows,
                       NCI_GRID, dx, inv_dx, d, n_fb, d_lg, static_cast<float *>(const_cast<void *>(ni.tab)));
    hipLaunchKernelGGL(nci_tail_kernel<float>, dim3(nbt), dim3(tb), 0, c->stream, d, rows, NCI_TAIL, NCI_TAIL_W0,
                       hw, d_lg, static_cast<float *>(const_cast<void *>(ni.tail)));
    HIPCHECK(launch_nci_rowmean<float>(static_cast<const float *>(ni.tab), rows, NCI_GRID, dx, d,
                                       const_cast<double *>(ni.mean), c->stream));
  } else {
    hipLaunchKernelGGL(nci_pack_kernel<double>, dim3(nb), dim3(tb), 0, c->stream, c->tmp.as<double>(), rows,
                       NCI_GRID, dx, inv_dx, d, n_fb, d_lg, static_cast<double *>(const_cast<void *>(ni.tab)));
    hipLaunchKernelGGL(nci_tail_kernel<double>, dim3(nbt), dim3(tb), 0, c->stream, d, rows, NCI_TAIL, NCI_TAIL_W0,
                       hw, d_lg, static_cast<double *>(const_cast<void *>(ni.tail)));
    HIPCHECK(launch_nci_rowmean<double>(static_cast<const double *>(ni.tab), rows, NCI_GRID, dx, d,
                                        const_cast<double *>(ni.mean), c->stream));
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));  // the scratch rows are reused by the next table
  c->nci[key] = std::move(buf);
  return ni;
}

// Under the reference's table semantics the rows above its row count and the
// upper tail u >= 0.999 are exact chi2 quantiles (VolSchemes.hpp:362-369).
// The engine evaluates them from the extended rows (exact slopes) for u in
// [NCI_U_ACC grid steps, NCI_U_ACC_HI), and from the tail rows for
// u >= 0.999, whose interpolation agrees with the exact inversion to ~1e-11
// and ~3e-11 relative there (the iterative inversion costs a whole
// wavefront several thousand cycles and fired on ~4 % of the Bates strip's
// steps); nearer u = 0 and in [NCI_U_ACC_HI, 0.999), where the uniform grid
// is coarse against the quantile's curvature, the rows above the
// reference's count are still inverted exactly.
constexpr int NCI_U_ACC = 100;
constexpr double NCI_U_ACC_HI = 0.99;

// Table fields of a job factor; the exact inversion of the first grid
// interval is used unless the config asks for the reference's table
// semantics.
template <typename R>
void set_nci(Job<R> &J, int f, const NciInfo &ni, bool reference) {
  J.nci[f] = ni.tab;
  J.nci_lgam[f] = ni.lg;
  J.nci_lg_n = NCI_LG;
  J.nci_grid = ni.grid;
  J.nci_dx = ni.dx;
  J.nci_inv_dx = ni.inv_dx;
  J.nci_tail[f] = ni.tail;
  J.nci_mean[f] = ni.mean;
  J.nci_rows[f] = ni.rows;
  J.nci_u_lo = reference ? 0.0 : ni.dx;
  J.nci_rows_acc[f] = reference ? ni.rows - 1 : -1;
  J.nci_u_acc = NCI_U_ACC * ni.dx;
  J.nci_u_acc_hi = NCI_U_ACC_HI;
  J.tail_pts = NCI_TAIL;
  J.tail_w0 = NCI_TAIL_W0;
  J.tail_inv_hw = (NCI_TAIL - 1) / (NCI_TAIL_W1 - NCI_TAIL_W0);
}

// ---------------------------------------------------------------------------
// job building
// ---------------------------------------------------------------------------
double expected_avg_variance(const myql_model &m, double T) {
  auto h = [&](const myql_heston_params &p) {
    if (std::fabs(p.kappa * T) > 1e-8) return p.theta + (p.v0 - p.theta) * (1.0 - std::exp(-p.kappa * T)) / (p.kappa * T);
    return p.v0;
  };
  double v = 0.0;
  if (m.n_factors == 0) v += m.vol * m.vol;
  else if (m.n_factors == 1) v += h(m.heston[0]);
  else v += h(m.heston[0]) + h(m.heston[1]);
  if (m.jump_kind == MYQL_JUMP_MERTON) v += m.jump[0] * (m.jump[1] * m.jump[1] + m.jump[2] * m.jump[2]);
  else if (m.jump_kind == MYQL_JUMP_KOU)
    v += m.jump[0] * (m.jump[1] / (m.jump[2] * m.jump[2]) + (1.0 - m.jump[1]) / (m.jump[3] * m.jump[3]));
  return v;
}

// ---------------------------------------------------------------------------
// Randomized quasi-Monte Carlo (MYQL_STREAMS_SOBOL)
// ---------------------------------------------------------------------------
uint64_t qmc_reps(const myql_mc_config &cfg) { return cfg.qmc_replicates ? cfg.qmc_replicates : 16; }

// Uniforms a path draws, at most: per step the stock's normal, one per NV
// factor and two per NCI factor (the Poisson and chi-squared uniforms), two
// per step for a path-dependent instrument's Merton jump (count, size), and
// two for a European's terminal Merton jump; a zero-factor European takes
// one step to maturity.
uint64_t qmc_dims(const myql_model &m, const int32_t *schemes, const myql_instrument &in, uint64_t steps) {
  const bool pd = in.kind != MYQL_INSTR_EUROPEAN;
  const bool jumps = m.jump_kind == MYQL_JUMP_MERTON;
  uint64_t per_step = 1;
  for (int f = 0; f < m.n_factors; ++f) per_step += schemes[f] == MYQL_SCHEME_NCI ? 2 : 1;
  if (pd && jumps) per_step += 2;
  const uint64_t n = (m.n_factors == 0 && !pd) ? 1 : steps;
  return n * per_step + (!pd && jumps ? 2 : 0);
}

uint64_t splitmix64_host(uint64_t &x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// x has order 2^s - 1 modulo p (bit i = coefficient of x^i, degree s).
bool gf2_primitive(uint32_t p, int s) {
  auto mulmod = [&](uint32_t a, uint32_t b) {
    uint32_t r = 0;
    while (b) {
      if (b & 1) r ^= a;
      b >>= 1;
      a <<= 1;
      if ((a >> s) & 1) a ^= p;
    }
    return r;
  };
  auto xpow = [&](uint32_t e) {
    uint32_t r = 1, base = 2;
    while (e) {
      if (e & 1) r = mulmod(r, base);
      base = mulmod(base, base);
      e >>= 1;
    }
    return r;
  };
  if (!((p >> s) & 1) || !(p & 1)) return false;
  const uint32_t order = (1u << s) - 1;
  if (xpow(order) != 1) return false;
  uint32_t n = order;
  for (uint32_t f = 2; f * f <= n; ++f) {
    if (n % f) continue;
    if (xpow(order / f) == 1) return false;
    while (n % f == 0) n /= f;
  }
  return n == 1 || xpow(order / n) != 1;
}

// Sobol' direction integers V[k][d] (k = 0..31 for bits 1..32 of the
// point index, d < QMC_DMAX), Bratley & Fox: coordinate 0 the van der
// Corput sequence; coordinate d >= 1 the d-th primitive polynomial over
// GF(2) in order of degree and encoding, odd initial direction integers
// m_k < 2^k from splitmix64 seeded with d, and the recurrence
// m_k = 2 a_1 m_(k-1) ^ ... ^ 2^(s-1) a_(s-1) m_(k-s+1) ^ 2^s m_(k-s) ^ m_(k-s).
std::vector<uint32_t> sobol_directions() {
  std::vector<uint32_t> V(32 * QMC_DMAX);
  for (int k = 1; k <= 32; ++k) V[(k - 1) * QMC_DMAX] = 1u << (32 - k);
  int d = 1;
  for (int s = 1; d < QMC_DMAX; ++s) {
    for (uint32_t mid = 0; mid < (1u << (s - 1)) && d < QMC_DMAX; ++mid) {
      const uint32_t p = (1u << s) | (mid << 1) | 1u;
      if (!gf2_primitive(p, s)) continue;
      uint64_t m[33] = {};
      uint64_t st = static_cast<uint64_t>(d);
      for (int k = 1; k <= s && k <= 32; ++k) {
        const uint64_t z = splitmix64_host(st);
        m[k] = k > 1 ? 2 * (z % (1ull << (k - 1))) + 1 : 1;
      }
      for (int k = s + 1; k <= 32; ++k) {
        uint64_t v = m[k - s] ^ (m[k - s] << s);
        for (int i = 1; i < s; ++i)
          if ((p >> (s - i)) & 1) v ^= m[k - i] << i;
        m[k] = v;
      }
      for (int k = 1; k <= 32; ++k) V[(k - 1) * QMC_DMAX + d] = static_cast<uint32_t>((m[k] << (32 - k)) & 0xFFFFFFFFull);
      ++d;
    }
  }
  return V;
}

// Direction and lower-bit tables on the context's device (built once).
void qmc_tables(myql_ctx *c) {
  if (c->qmc_V.p) return;
  const std::vector<uint32_t> V = sobol_directions();
  std::vector<uint32_t> lo(static_cast<size_t>(QMC_DMAX) * 64);
  for (int d = 0; d < QMC_DMAX; ++d)
    for (int l = 0; l < 64; ++l) {
      uint32_t x = 0;
      for (int k = 0; k < 6; ++k)
        if ((l >> k) & 1) x ^= V[k * QMC_DMAX + d];
      lo[static_cast<size_t>(d) * 64 + l] = x;
    }
  c->qmc_V.reserve(V.size() * sizeof(uint32_t));
  c->qmc_lo.reserve(lo.size() * sizeof(uint32_t));
  HIPCHECK(hipMemcpyAsync(c->qmc_V.p, V.data(), V.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->qmc_lo.p, lo.data(), lo.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

// Digital shift of replicate `rep`: splitmix64 seeded with
// seed ^ rep * 0x9E3779B97F4A7C15, the low 32 bits of one output per
// coordinate.
void qmc_shift(uint64_t seed, uint64_t rep, int dims, uint32_t *out) {
  uint64_t st = seed ^ (rep * 0x9E3779B97F4A7C15ull);
  for (int d = 0; d < dims; ++d) out[d] = static_cast<uint32_t>(splitmix64_host(st) & 0xFFFFFFFFull);
}

void validate(const myql_model *m, const int32_t *schemes, const myql_instrument *in, const myql_mc_config *cfg) {
  if (!m || !schemes || !in || !cfg) throw Fail(MYQL_E_INVALID_ARG, "null argument");
  if (m->n_factors < 0 || m->n_factors > 2) throw Fail(MYQL_E_INVALID_ARG, "n_factors must be 0, 1 or 2");
  for (int i = 0; i < m->n_factors; ++i)
    if (schemes[i] < MYQL_SCHEME_NV || schemes[i] > MYQL_SCHEME_EXACT)
      throw Fail(MYQL_E_INVALID_ARG, "scheme must be NV, NCI or Exact for each variance factor");
  if (m->jump_kind < MYQL_JUMP_NONE || m->jump_kind > MYQL_JUMP_KOU) throw Fail(MYQL_E_INVALID_ARG, "bad jump kind");
  if (in->n_strikes < 1 || !in->strikes) throw Fail(MYQL_E_INVALID_ARG, "instrument needs strikes");
  if (in->kind < MYQL_INSTR_EUROPEAN || in->kind > MYQL_INSTR_BARRIER) throw Fail(MYQL_E_INVALID_ARG, "bad instrument kind");
  if (!(in->maturity > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "maturity must be positive");
  if (cfg->num_paths < 1 || cfg->time_steps < 1) throw Fail(MYQL_E_INVALID_ARG, "num_paths and time_steps must be >= 1");
  if (cfg->time_steps > (1ull << 30)) throw Fail(MYQL_E_INVALID_ARG, "time_steps too large");
  if (cfg->precision != MYQL_FP64 && cfg->precision != MYQL_FP32) throw Fail(MYQL_E_INVALID_ARG, "bad precision");
  if (cfg->stream_layout < MYQL_STREAMS_JUMP || cfg->stream_layout > MYQL_STREAMS_SOBOL)
    throw Fail(MYQL_E_INVALID_ARG, "bad stream layout");
  if (cfg->stream_layout == MYQL_STREAMS_SOBOL) {
    for (int i = 0; i < m->n_factors; ++i)
      if (schemes[i] == MYQL_SCHEME_EXACT)
        throw Fail(MYQL_E_UNSUPPORTED, "quasi-Monte Carlo: SchemeExact draws an unbounded number of uniforms");
    if (m->jump_kind == MYQL_JUMP_KOU)
      throw Fail(MYQL_E_UNSUPPORTED, "quasi-Monte Carlo: Kou jumps draw an unbounded number of uniforms");
    if (qmc_dims(*m, schemes, *in, cfg->time_steps) > QMC_DMAX)
      throw Fail(MYQL_E_UNSUPPORTED, "quasi-Monte Carlo: more than 1,024 coordinates per path");
    const uint64_t R = qmc_reps(*cfg);
    if (R < 2) throw Fail(MYQL_E_INVALID_ARG, "quasi-Monte Carlo: the standard error needs at least 2 replicates");
    if (cfg->num_paths < R) throw Fail(MYQL_E_INVALID_ARG, "quasi-Monte Carlo: fewer paths than replicates");
    if (cfg->num_paths / R > (1ull << 30)) throw Fail(MYQL_E_INVALID_ARG, "quasi-Monte Carlo: over 2^30 points per replicate");
  }
}

// Stepper constants of one variant (ASVJstepper.hpp:85-117, init_cir_impl
// :262-283, Scheme::prepare, JumpPolicy::prepare).
template <typename R>
void fill_variant(VarParams<R> &V, const myql_model &m, const int32_t *schemes, double dt, double r, double q,
                  double total_time, uint64_t steps, bool pathdep, const int nmax[2]) {
  std::memset(&V, 0, sizeof(V));
  V.dt = static_cast<R>(dt);
  V.total_time = static_cast<R>(static_cast<double>(steps) * dt);
  if (m.n_factors == 0) {
    const double var = m.vol * m.vol;
    const double gd = (r - q - 0.5 * var) * dt, gv = m.vol * std::sqrt(dt);
    V.gdrift = static_cast<R>(gd);
    V.gvol = static_cast<R>(gv);
    V.gdrift_n = static_cast<R>(gd * static_cast<double>(steps));
    V.gvol_n = static_cast<R>(gv * std::sqrt(static_cast<double>(steps)));
    V.gdrift_nd = gd * static_cast<double>(steps);
    V.gvol_nd = gv * std::sqrt(static_cast<double>(steps));
  } else {
    const double r_part = (r - q) / static_cast<double>(m.n_factors);
    for (int f = 0; f < m.n_factors; ++f) {
      const myql_heston_params &p = m.heston[f];
      const double h = dt;
      const double C0 = (r_part - p.rho * p.kappa * p.theta / p.sigma) * h;
      const double term = (p.kappa * p.rho / p.sigma - 0.5) * h * 0.5;
      V.C[f][0] = static_cast<R>(C0);
      V.C[f][1] = static_cast<R>(term - p.rho / p.sigma);
      V.C[f][2] = static_cast<R>(term + p.rho / p.sigma);
      V.C[f][3] = static_cast<R>(h * (1.0 - p.rho * p.rho) * 0.5);
      V.Cd[f][0] = C0;
      V.Cd[f][1] = term - p.rho / p.sigma;
      V.Cd[f][2] = term + p.rho / p.sigma;
      V.Cd[f][3] = h * (1.0 - p.rho * p.rho) * 0.5;
      V.v0[f] = static_cast<R>(p.v0);
      V.v0d[f] = p.v0;
      double K[3] = {0, 0, 0};
      if (schemes[f] == MYQL_SCHEME_NV) {
        if (p.sigma * p.sigma > 4.0 * p.kappa * p.theta)
          throw Fail(MYQL_E_NV_FELLER, "SchemeNV violation: sigma^2 > 4*kappa*theta");
        const double ex = std::exp(-p.kappa * h * 0.5);
        const double psi = (std::fabs(p.kappa) < 1e-10) ? h * 0.5 : (1.0 - ex) / p.kappa;
        K[0] = ex;
        K[1] = (p.kappa * p.theta - p.sigma * p.sigma * 0.25) * psi;
        K[2] = p.sigma * std::sqrt(h) * 0.5;
      } else {
        const double ex = std::exp(-p.kappa * h);
        const double psi = (std::fabs(p.kappa) < 1e-10) ? h : (1.0 - ex) / p.kappa;
        K[0] = (2.0 / (p.sigma * p.sigma * psi)) * ex;
        if (schemes[f] == MYQL_SCHEME_NCI) {
          K[1] = 4.0 * p.kappa * p.theta / (p.sigma * p.sigma);
          K[2] = 0.25 * p.sigma * p.sigma * psi;
        } else {
          K[1] = 2.0 * p.kappa * p.theta / (p.sigma * p.sigma);
          K[2] = 0.5 * p.sigma * p.sigma * psi;
        }
      }
      for (int i = 0; i < 3; ++i) {
        V.Kd[f][i] = K[i];
        V.Kr[f][i] = static_cast<R>(K[i]);
      }
      V.nmax[f] = nmax[f];
    }
  }
  const double span = pathdep ? dt : total_time;
  if (m.jump_kind == MYQL_JUMP_MERTON) {
    const double lam = m.jump[0], mu = m.jump[1], dl = m.jump[2];
    V.jmean = lam * span;
    V.jexpm = std::exp(-V.jmean);
    V.jdrift_d = -lam * (std::exp(mu + 0.5 * dl * dl) - 1.0) * span;
    V.jdrift = static_cast<R>(V.jdrift_d);
  } else if (m.jump_kind == MYQL_JUMP_KOU) {
    const double lam = m.jump[0], p = m.jump[1], e1 = m.jump[2], e2 = m.jump[3];
    const double k = (p * e1) / (e1 - 1.0) + ((1.0 - p) * e2) / (e2 + 1.0) - 1.0;
    V.jmean = lam * span;
    V.jexpm = std::exp(-V.jmean);
    V.jdrift_d = -lam * k * span;
    V.jdrift = static_cast<R>(V.jdrift_d);
  } else {
    V.jmean = 0.0;
    V.jexpm = 1.0;
  }
  // The reference draws jump counts with Boost's poisson_distribution
  // (JumpPolicies.hpp:38,48,86,98), which inverts one uniform for means
  // below 10 (restated here) and switches to transformed rejection (PTRD,
  // a variable number of draws) from 10 on; from 10 on the kernels invert
  // the same single uniform exactly instead (jump_count): the same
  // distribution from a different draw sequence.
}

struct JobSpec {
  const myql_model *model;
  const int32_t *schemes;
  const myql_instrument *instr;
  int mode;
  const myql_mc_config *cfg;
  double S0, r, q;
  myql_mc_result *res;
  int k0, nk;  // strike chunk [k0, k0+nk) of the instrument
  int qmc_rep = -1;  // MYQL_STREAMS_SOBOL: the replicate (digital shift) this job prices
};

// True when the two calls simulate identical paths: same model, schemes,
// config (paths, steps, seed, bumps, precision, streams, table rows), spot,
// rates, maturity and Greek mode, neither in flight. Their instruments may
// differ; the paths do not depend on them.
bool same_paths(const JobSpec &x, const JobSpec &y) {
  const myql_mc_config &a = *x.cfg, &b = *y.cfg;
  const int nf = x.model->n_factors;
  if (x.qmc_rep != y.qmc_rep) return false;
  if (std::memcmp(x.model, y.model, sizeof(myql_model)) != 0) return false;
  for (int f = 0; f < nf; ++f)
    if (x.schemes[f] != y.schemes[f]) return false;
  const uint32_t keep = ~static_cast<uint32_t>(MYQL_FLAG_AUTOTUNE);
  return a.num_paths == b.num_paths && a.time_steps == b.time_steps && a.seed == b.seed && a.fd_bump == b.fd_bump &&
         a.vol_bump == b.vol_bump && a.t_bump == b.t_bump && a.r_bump == b.r_bump && a.precision == b.precision &&
         a.stream_layout == b.stream_layout && a.num_streams == b.num_streams && (a.flags & keep) == (b.flags & keep) &&
         x.S0 == y.S0 && x.r == y.r && x.q == y.q && x.mode == y.mode && x.instr->maturity == y.instr->maturity &&
         x.instr->t_elapsed == 0.0 && y.instr->t_elapsed == 0.0;
}

struct BuiltJob {
  int nf, nv;        // kernel instantiation
  bool pd;
  bool qmc;          // randomized quasi-Monte Carlo points
  int n_slots;       // work units
  uint64_t cost;     // per-unit cost estimate
  int nk;
  JobSpec spec;
  double df;
  uint64_t variants_alg;
};

int pow2ceil(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Model, instrument and tracker fields of a job (everything but the
// variants): get_tracker_config<Mode> and the payoff, with the automatic
// digital smoothing of MonteCarloPricer.hpp:51-74 when `auto_smooth`.
template <typename R>
void fill_instrument(Job<R> &J, const myql_model &M0, const int32_t *schemes, const myql_instrument &in, int mode,
                     double S0, double r, double fd_bump, bool auto_smooth) {
  const bool ess = mode != MYQL_GREEK_NONE;
  const int nf = M0.n_factors;
  const bool pd = in.kind != MYQL_INSTR_EUROPEAN;
  const double T = in.maturity;
  const double h = S0 * fd_bump;
  (void)r;
  std::memset(&J, 0, sizeof(J));
  J.nf = nf;
  J.scheme[0] = nf > 0 ? schemes[0] : 0;
  J.scheme[1] = nf > 1 ? schemes[1] : 0;
  J.jump_kind = M0.jump_kind;
  J.pathdep = pd;
  J.jmu = static_cast<R>(M0.jump[1]);
  J.jdelta = static_cast<R>(M0.jump[2]);
  J.kp = static_cast<R>(M0.jump[1]);
  J.keta1 = static_cast<R>(M0.jump[2]);
  J.keta2 = static_cast<R>(M0.jump[3]);
  J.tracker = in.kind;
  J.payoff = in.payoff;
  J.call = in.option_type == MYQL_CALL;
  J.fixed_strike = in.fixed_strike;
  J.barrier_up = in.barrier_dir == MYQL_BARRIER_UP;
  J.barrier_in = in.barrier_action == MYQL_BARRIER_IN;
  J.rebate_timing = in.rebate_timing;
  J.mode = mode;
  J.S0 = S0;
  J.logS0 = std::log(S0);
  J.mult_up = ess ? (S0 + h) / S0 : 1.0;
  J.mult_dn = ess ? (S0 - h) / S0 : 1.0;
  J.inv2h = 1.0 / (2.0 * h);
  J.invh2 = 1.0 / (h * h);
  if (in.kind == MYQL_INSTR_BARRIER) {
    const double bb = std::log(in.barrier);
    J.b_base = static_cast<R>(bb);
    J.b_up = static_cast<R>(ess ? std::log(in.barrier * S0 / (S0 + h)) : bb);
    J.b_dn = static_cast<R>(ess ? std::log(in.barrier * S0 / (S0 - h)) : bb);
  }
  J.init_sumlog = in.past_avg_log * in.t_elapsed;
  J.init_sumS = in.past_avg_S * in.t_elapsed;
  J.hist_min_logS = in.hist_min_S > 0.0 ? std::log(in.hist_min_S) : 1e300;
  J.hist_max_logS = in.hist_max_S > 0.0 ? std::log(in.hist_max_S) : -1e300;
  J.t_elapsed = in.t_elapsed;
  J.maturity = T;
  J.rebate = in.rebate;
  J.rebate_rate = in.rebate_rate;
  // automatic digital smoothing (MonteCarloPricer.hpp:51-74)
  double eps = in.payoff_eps;
  if (auto_smooth && ess && in.payoff != MYQL_PAYOFF_VANILLA && eps < 0.0) {
    const double sig = std::sqrt(expected_avg_variance(M0, T));
    const double scale = in.is_strip ? S0 : in.strikes[0];
    eps = sig * scale * std::sqrt(T) / 10.0;
  }
  J.eps_scale = 0.70710678118654752440 / eps;
  J.n_fields = in.kind == MYQL_INSTR_EUROPEAN ? 1
               : in.kind != MYQL_INSTR_BARRIER ? 2
               : (in.rebate_timing == MYQL_REBATE_AT_HIT ? NFIELD : 3);
}

template <typename R>
BuiltJob build_job(myql_ctx *c, const JobSpec &js, Job<R> &J, std::vector<double> &strike_store) {
  const myql_model &M0 = *js.model;
  const myql_instrument &in = *js.instr;
  const myql_mc_config &cfg = *js.cfg;
  const int mode = js.mode;
  const bool full = mode == MYQL_GREEK_FULL;
  const int nf = M0.n_factors;
  const bool pd = in.kind != MYQL_INSTR_EUROPEAN;
  const double S0 = js.S0, r = js.r, q = js.q;
  const double T = in.maturity;
  const uint64_t steps = cfg.time_steps, M = cfg.num_paths;
  const double dt = T / static_cast<double>(steps);

  fill_instrument(J, M0, js.schemes, in, mode, S0, r, cfg.fd_bump, true);

  // Greek bump constants (MonteCarloPricer.hpp:297-300,384-399)
  J.inv2dv = 1.0 / (2.0 * cfg.vol_bump);
  J.inv2dT = 1.0 / (2.0 * cfg.t_bump);
  J.inv2dr = 1.0 / (2.0 * cfg.r_bump);
  J.cv[0] = nf > 0 ? 2.0 * std::sqrt(M0.heston[0].v0) : 1.0;
  J.cv[1] = nf > 1 ? 2.0 * std::sqrt(M0.heston[1].v0) : 1.0;
  J.dfTu = std::exp(-r * cfg.t_bump);
  J.dfTd = std::exp(r * cfg.t_bump);
  J.dfru = std::exp(-cfg.r_bump * T);
  J.dfrd = std::exp(cfg.r_bump * T);
  J.erb_up = std::exp(cfg.r_bump * T);
  J.erb_dn = std::exp(-cfg.r_bump * T);
  J.nvega = std::max(nf, 1);

  // variants: models, step sizes, rates, maturities
  struct VDef {
    int id;
    myql_model m;
    double dt, r, T;
  };
  std::vector<VDef> vd;
  vd.push_back({V_BASE, M0, dt, r, T});
  if (full) {
    for (int f = 0; f < J.nvega; ++f) {
      myql_model mu = M0, md = M0;
      if (nf == 0) {
        mu.vol += cfg.vol_bump;
        md.vol -= cfg.vol_bump;
      } else {
        mu.heston[f].v0 += cfg.vol_bump;
        md.heston[f].v0 -= cfg.vol_bump;
      }
      vd.push_back({V_V0UP + 2 * f, mu, dt, r, T});
      vd.push_back({V_V0DN + 2 * f, md, dt, r, T});
    }
    const double Tu = T + cfg.t_bump, Td = T - cfg.t_bump;
    vd.push_back({V_TUP, M0, Tu / steps, r, Tu});
    vd.push_back({V_TDN, M0, Td / steps, r, Td});
    if (pd) {
      vd.push_back({V_RUP, M0, dt, r + cfg.r_bump, T});
      vd.push_back({V_RDN, M0, dt, r - cfg.r_bump, T});
    }
  }
  J.derive_r = full && !pd;

  // NCI tables: one per factor, rows for the largest workspace among
  // variants (under the reference's semantics the variants' own row counts
  // are the reference's, and the table still holds the extended rows that
  // serve its exact fallback)
  const bool ref_rows = cfg.flags & MYQL_FLAG_REFERENCE_NCI_ROWS;
  int nmax_max[2] = {0, 0};
  std::vector<std::array<int, 2>> nmaxv(vd.size(), {0, 0});
  for (size_t i = 0; i < vd.size(); ++i)
    for (int f = 0; f < nf; ++f)
      if (js.schemes[f] == MYQL_SCHEME_NCI) {
        nmaxv[i][f] = nci_nmax(vd[i].m.heston[f], vd[i].dt, ref_rows);
        nmax_max[f] = std::max(nmax_max[f], nci_nmax(vd[i].m.heston[f], vd[i].dt, false));
      }
  for (int f = 0; f < nf; ++f)
    if (js.schemes[f] == MYQL_SCHEME_NCI) {
      if (c->noalloc)
        throw Fail(MYQL_E_UNSUPPORTED, "myql_mc_calculate_noalloc: SchemeNCI needs its chi-squared table "
                                       "allocated (use myql_mc_calculate)");
      const myql_heston_params &p = M0.heston[f];
      const double d = 4.0 * p.kappa * p.theta / (p.sigma * p.sigma);
      NciInfo ni = nci_table(c, d, nmax_max[f], cfg.precision, nci_nmax(p, dt, true));
      set_nci(J, f, ni, ref_rows);
      if (cfg.precision == MYQL_FP32 && full) {
        // FP64 rows for the maturity-bumped variants stepped in FP64
        const NciInfo n64 = nci_table(c, d, nmax_max[f], MYQL_FP64, nci_nmax(p, dt, true));
        J.nci64[f] = static_cast<const double *>(n64.tab);
        J.nci_tail64[f] = static_cast<const double *>(n64.tail);
      }
    }
  for (size_t i = 0; i < vd.size(); ++i)
    fill_variant(J.var[vd[i].id], vd[i].m, js.schemes, vd[i].dt, vd[i].r, q, vd[i].T, steps, pd,
                 nmaxv[i].data());

  // grouping: lockstep when every variant consumes the same draws
  bool lockstep = true;
  for (int f = 0; f < nf; ++f)
    if (js.schemes[f] == MYQL_SCHEME_EXACT) lockstep = false;
  if (pd && M0.jump_kind != MYQL_JUMP_NONE) lockstep = false;
  const int nsim = static_cast<int>(vd.size());
  for (int i = 0; i < nsim; ++i) {
    J.var_list[i] = vd[i].id;
    J.staged[vd[i].id] = 1;
  }
  int nv;
  if (lockstep && nsim > 1) {
    nv = nsim;
    J.n_groups = 1;
    J.group_first[0] = 0;
    J.group_size[0] = nsim;
  } else {
    nv = 1;
    J.n_groups = nsim;
    for (int i = 0; i < nsim; ++i) {
      J.group_first[i] = i;
      J.group_size[i] = 1;
    }
  }
  if (!mc_supported(nf, nv, pd)) throw Fail(MYQL_E_UNSUPPORTED, "no kernel for this factor count / variant set");
  uint64_t S = cfg.num_streams ? cfg.num_streams : std::min<uint64_t>(65536, M);
  if (S > M) S = M;
  J.M = M;
  J.S = S;
  J.steps = steps;
  if (cfg.stream_layout == MYQL_STREAMS_SOBOL) {
    // the replicate's shift is placed by run_jobs (qmc_shift: an offset
    // into the call's shift buffer until then)
    qmc_tables(c);
    J.streams = nullptr;
    J.qmc_V = c->qmc_V.as<uint32_t>();
    J.qmc_lo = c->qmc_lo.as<uint32_t>();
    J.qmc_dims = static_cast<int>(qmc_dims(M0, js.schemes, in, steps));
  } else {
    J.streams = stream_states(c, cfg.seed, cfg.stream_layout, S);
  }

  // strikes of this chunk
  J.nk = js.nk;
  J.nkp = pow2ceil(js.nk);
  const size_t off = strike_store.size();
  for (int k = 0; k < js.nk; ++k) strike_store.push_back(in.strikes[js.k0 + k]);
  J.strikes = reinterpret_cast<const double *>(off);  // rebased after upload

  BuiltJob b;
  b.nf = nf;
  b.nv = nv;
  b.pd = pd;
  b.qmc = cfg.stream_layout == MYQL_STREAMS_SOBOL;
  b.n_slots = static_cast<int>((S + BLOCK - 1) / BLOCK);
  const uint64_t per_stream = (M + S - 1) / S;
  b.cost = per_stream * steps * static_cast<uint64_t>(nsim) * (nf == 0 ? 1 : nf) + per_stream * js.nk;
  b.nk = js.nk;
  b.spec = js;
  b.df = std::exp(-r * T);
  b.variants_alg = full ? (nf == 2 ? 9 : 7) : 1;
  return b;
}

// Run a set of jobs of one path precision. Jobs with the same kernel
// instantiation share one persistent launch.
template <typename R>
void run_jobs(myql_ctx *c, const std::vector<JobSpec> &specs, double &device_ms) {
  std::vector<Job<R>> jobs(specs.size());
  std::vector<BuiltJob> built;
  std::vector<double> strikes;
  for (size_t i = 0; i < specs.size(); ++i) built.push_back(build_job<R>(c, specs[i], jobs[i], strikes));

  // Fused instruments: an arithmetic Asian whose paths equal a barrier
  // call's is priced from that call's staged samples (the simulation keeps
  // both trackers' state) instead of simulating the same paths again.
  std::vector<char> is_view(jobs.size(), 0);
  {
    std::vector<char> taken(jobs.size(), 0);
    for (size_t b = 0; b < jobs.size(); ++b) {
      if (taken[b] || built[b].spec.instr->kind != MYQL_INSTR_BARRIER) continue;
      for (size_t a = 0; a < jobs.size(); ++a) {
        if (a == b || taken[a] || built[a].spec.instr->kind != MYQL_INSTR_ARITH_ASIAN) continue;
        if (built[a].nf != built[b].nf || built[a].nv != built[b].nv || built[a].pd != built[b].pd ||
            !same_paths(built[a].spec, built[b].spec))
          continue;
        taken[a] = taken[b] = 1;
        is_view[a] = 1;
        jobs[b].fuse_asian = 1;
        jobs[b].view = static_cast<int>(a) + 1;
        jobs[b].n_fields = std::max(jobs[b].n_fields, jobs[a].n_fields);
        break;
      }
    }
  }

  // digital shifts of the quasi-Monte Carlo jobs' replicates
  {
    size_t words = 0;
    for (size_t i = 0; i < jobs.size(); ++i)
      if (built[i].qmc) words += static_cast<size_t>(jobs[i].qmc_dims);
    if (words) {
      std::vector<uint32_t> sh(words);
      size_t off = 0;
      std::vector<size_t> at(jobs.size(), 0);
      for (size_t i = 0; i < jobs.size(); ++i) {
        if (!built[i].qmc) continue;
        at[i] = off;
        qmc_shift(built[i].spec.cfg->seed, static_cast<uint64_t>(built[i].spec.qmc_rep), jobs[i].qmc_dims,
                  sh.data() + off);
        off += static_cast<size_t>(jobs[i].qmc_dims);
      }
      uint32_t *d_sh = static_cast<uint32_t *>(dev_mem(c, c->qmc_shift, words * sizeof(uint32_t)));
      HIPCHECK(hipMemcpyAsync(d_sh, sh.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      for (size_t i = 0; i < jobs.size(); ++i)
        if (built[i].qmc) jobs[i].qmc_shift = d_sh + at[i];
    }
  }

  // kernel groups (one persistent launch each) and their launch shapes:
  // the tuned shape where the autotuner has chosen one, else the default
  using GKey = std::tuple<int, int, bool, bool>;  // nf, nv, pd, qmc
  const int prec = sizeof(R) == 4 ? MYQL_FP32 : MYQL_FP64;
  bool autotune = false;
  for (const BuiltJob &b : built) autotune = autotune || (b.spec.cfg->flags & MYQL_FLAG_AUTOTUNE);
  if (c->noalloc) autotune = false;  // trial launches would need buffers of their own
  auto shape_of = [&](const GKey &k) -> std::pair<McShape, int> {
    if (std::get<3>(k)) return {MC_QMC_SHAPE, 4};
    auto it = c->tuned.find(std::make_tuple(prec, std::get<0>(k), std::get<1>(k), std::get<2>(k)));
    if (it != c->tuned.end()) return {it->second.shape, it->second.gridf};
    return {MC_DEFAULT_SHAPE, 4};
  };
  auto gkey = [&](size_t i) { return GKey{built[i].nf, built[i].nv, built[i].pd, built[i].qmc}; };
  auto slots_for = [&](size_t i, int blk) { return static_cast<int>((jobs[i].S + blk - 1) / blk); };

  // slots and sums layout; partials are laid out for the smallest block so
  // that any shape (trial or chosen) fits
  int soff = 0;
  std::vector<int> n_slots(jobs.size()), nks(jobs.size()), out_off(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) {
    n_slots[i] = slots_for(i, shape_of(gkey(i)).first.blk);
    nks[i] = built[i].nk;
    out_off[i] = soff;
    soff += 2 * NSTAT * built[i].nk;
  }
  int min_blk = BLOCK;
  for (const McShape &sh : MC_TUNE_SHAPES) min_blk = std::min(min_blk, sh.blk);
  std::vector<size_t> pbase(jobs.size());
  size_t partial_elems = 0;
  uint64_t stream_total = 0;
  for (size_t i = 0; i < jobs.size(); ++i) {
    pbase[i] = partial_elems;
    partial_elems += static_cast<size_t>(slots_for(i, min_blk)) * 2 * NSTAT * nks[i];
    stream_total += jobs[i].S + BLOCK;
  }
  double *d_partial = static_cast<double *>(dev_mem(c, c->partial, std::max<size_t>(partial_elems, 1) * sizeof(double)));
  double *d_sums = static_cast<double *>(dev_mem(c, c->sums, std::max(soff, 1) * sizeof(double)));
  const size_t strike_bytes = strikes.size() * sizeof(double);
  const size_t jobs_bytes = jobs.size() * sizeof(Job<R>);
  char *d_jobs_raw = static_cast<char *>(dev_mem(c, c->jobs, jobs_bytes + strike_bytes + 256));
  double *d_strikes = reinterpret_cast<double *>(d_jobs_raw + ((jobs_bytes + 255) / 256) * 256);
  double *d_vaccg = nullptr;
  if (std::find(is_view.begin(), is_view.end(), 1) != is_view.end())
    d_vaccg = static_cast<double *>(dev_mem(c, c->vaccg, static_cast<size_t>(stream_total) * 2 * NSTAT * sizeof(double)));
  for (size_t i = 0; i < jobs.size(); ++i) {
    if (jobs[i].view) jobs[i].view_accg = d_vaccg;
    jobs[i].strikes = d_strikes + reinterpret_cast<size_t>(jobs[i].strikes);
    // each job indexes its partials from its own base with first_slot = 0
    jobs[i].partial = d_partial + pbase[i];
    jobs[i].first_slot = 0;
  }
  // per-unit accumulators: 2*NSTAT*blk doubles per unit of blk streams
  double *d_accg = static_cast<double *>(dev_mem(c, c->accg, static_cast<size_t>(stream_total) * 2 * NSTAT * sizeof(double)));

  std::map<GKey, std::vector<size_t>> members;  // jobs with work units (views have none)
  for (size_t i = 0; i < jobs.size(); ++i)
    if (!is_view[i]) members[gkey(i)].push_back(i);
  // work units, queue counters and reduction meta; sized for the smallest
  // block, so any launch shape fits (every buffer of the call is in place
  // before the first copy)
  size_t max_units = 0;
  for (size_t i = 0; i < jobs.size(); ++i)
    if (!is_view[i]) max_units += static_cast<size_t>(slots_for(i, min_blk));
  const size_t nj = jobs.size();
  Unit *d_units = static_cast<Unit *>(dev_mem(c, c->units, max_units * sizeof(Unit) + 64));
  unsigned *d_counter = static_cast<unsigned *>(dev_mem(c, c->counter, members.size() * 64));
  long long *d_base = static_cast<long long *>(dev_mem(c, c->meta, nj * (sizeof(long long) + 3 * sizeof(int)) + 64));
  HIPCHECK(hipMemcpyAsync(d_jobs_raw, jobs.data(), jobs_bytes, hipMemcpyHostToDevice, c->stream));
  if (strike_bytes)
    HIPCHECK(hipMemcpyAsync(d_strikes, strikes.data(), strike_bytes, hipMemcpyHostToDevice, c->stream));
  const Job<R> *d_jobs = reinterpret_cast<const Job<R> *>(d_jobs_raw);
  // units of one group for block size blk, most expensive job first
  auto units_for = [&](const std::vector<size_t> &idx, int blk) {
    std::vector<Unit> u;
    for (size_t i : idx)
      for (int b = 0; b < slots_for(i, blk); ++b) u.push_back({static_cast<int>(i), b});
    std::stable_sort(u.begin(), u.end(), [&](const Unit &a, const Unit &b) { return built[a.job].cost > built[b.job].cost; });
    return u;
  };
  auto staging = [&](const GKey &k, int blk) {
    int nfields = 1;
    for (size_t i : members[k]) nfields = std::max(nfields, jobs[i].n_fields);
    const size_t smem = static_cast<size_t>(nfields) * MAXV * blk * sizeof(R);
    if (smem > c->lds_max) throw Fail(MYQL_E_UNSUPPORTED, "staging area exceeds the device's LDS per block");
    return smem;
  };
  // Persistent grid: gridf x the resident blocks per CU (the occupancy
  // query, but never fewer than the VGPR budget allows: 512 per SIMD lane,
  // blk/64 waves per block, 4 SIMDs); surplus blocks find the queue empty.
  auto grid_for = [&](const GKey &k, McShape sh, int gridf, size_t smem, int n_units) {
    const int bpc = std::max(mc_blocks_per_cu<R>(std::get<0>(k), std::get<1>(k), std::get<2>(k), sh, smem),
                             2 * BLOCK / sh.blk);
    return std::max(1, std::min(n_units, gridf * bpc * c->n_cu));
  };

  // Autotuner: time every shape on trial launches of this call's own jobs
  // with an eighth of each stream's samples, keep the fastest.
  if (autotune) {
    for (auto &kv : members) {
      const GKey &k = kv.first;
      if (std::get<3>(k)) continue;  // one compiled shape
      const auto tkey = std::make_tuple(prec, std::get<0>(k), std::get<1>(k), std::get<2>(k));
      if (c->tuned.count(tkey)) continue;
      const int n_shapes = mc_tune_shape_count(std::get<0>(k), std::get<1>(k), std::get<2>(k));
      std::vector<Job<R>> tj = jobs;
      for (Job<R> &j : tj) j.M = j.S * std::max<uint64_t>(1, (j.M / j.S) / 8);
      c->tjobs.reserve(jobs_bytes);
      HIPCHECK(hipMemcpyAsync(c->tjobs.p, tj.data(), jobs_bytes, hipMemcpyHostToDevice, c->stream));
      c->tcounter.reserve(64);
      std::string log;
      float best = 0.f;
      McShape best_sh = MC_DEFAULT_SHAPE;
      int best_gf = 4;
      for (int si = 0; si < n_shapes; ++si) {
        const McShape sh = MC_TUNE_SHAPES[si];
        const std::vector<Unit> u = units_for(kv.second, sh.blk);
        c->tunits.reserve(u.size() * sizeof(Unit));
        HIPCHECK(hipMemcpyAsync(c->tunits.p, u.data(), u.size() * sizeof(Unit), hipMemcpyHostToDevice, c->stream));
        const size_t smem = staging(k, sh.blk);
        const int n_units = static_cast<int>(u.size());
        // rep 0 warms the kernel up, untimed; then grid factors 4 and 1
        // twice each, keeping each one's faster time (trials vary by ~10 %)
        float t_gf[2] = {0.f, 0.f};
        for (int rep = 0; rep < 5; ++rep) {
          const int gridf = (rep & 1) ? 4 : 1;
          HIPCHECK(hipMemsetAsync(c->tcounter.p, 0, 64, c->stream));
          HIPCHECK(hipEventRecord(c->ev0, c->stream));
          HIPCHECK(launch_mc<R>(std::get<0>(k), std::get<1>(k), std::get<2>(k), sh, c->tjobs.as<Job<R>>(),
                                c->tunits.as<Unit>(), n_units, c->tcounter.as<unsigned>(), d_accg,
                                smem, grid_for(k, sh, gridf, smem, n_units), c->stream));
          HIPCHECK(hipEventRecord(c->ev1, c->stream));
          HIPCHECK(hipEventSynchronize(c->ev1));
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
          if (rep == 0) continue;
          float &t = t_gf[gridf == 4 ? 0 : 1];
          t = (t == 0.f || ms < t) ? ms : t;
        }
        for (int gi = 0; gi < 2; ++gi) {
          const int gridf = gi == 0 ? 4 : 1;
          const float ms = t_gf[gi];
          char line[96];
          std::snprintf(line, sizeof line, "%s%d/%d/%dx%s %.3f ms", log.empty() ? "" : ", ", sh.blk, sh.unr, gridf,
                        sh.occ ? "/w4" : "", static_cast<double>(ms));
          log += line;
          if (best == 0.f || ms < best) {
            best = ms;
            best_sh = sh;
            best_gf = gridf;
          }
        }
      }
      c->tuned[tkey] = {best_sh, best_gf, log};
    }
    for (size_t i = 0; i < jobs.size(); ++i) n_slots[i] = slots_for(i, shape_of(gkey(i)).first.blk);
  }

  // units grouped by kernel instantiation, most expensive first
  std::vector<Unit> all_units;
  std::vector<std::pair<size_t, size_t>> ranges;
  std::vector<GKey> order;
  for (auto &kv : members) {
    const std::vector<Unit> u = units_for(kv.second, shape_of(kv.first).first.blk);
    ranges.push_back({all_units.size(), u.size()});
    all_units.insert(all_units.end(), u.begin(), u.end());
    order.push_back(kv.first);
  }
  if (all_units.size() > max_units) throw Fail(MYQL_E_HIP, "work-unit count exceeds its buffer");
  // queue counters start at zero (a fill kernel: callers may capture this
  // sequence in a graph, where memset nodes are avoided)
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, c->stream, d_counter,
                     static_cast<int>(members.size() * 16));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(d_units, all_units.data(), all_units.size() * sizeof(Unit), hipMemcpyHostToDevice,
                          c->stream));

  // meta for the reduction: per job the element offset of its partials,
  // its unit count, strike count and output offset
  std::vector<long long> mbase(pbase.begin(), pbase.end());
  std::vector<int> mint;
  mint.insert(mint.end(), n_slots.begin(), n_slots.end());
  mint.insert(mint.end(), nks.begin(), nks.end());
  mint.insert(mint.end(), out_off.begin(), out_off.end());
  int *d_mint = reinterpret_cast<int *>(d_base + nj);
  HIPCHECK(hipMemcpyAsync(d_base, mbase.data(), nj * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(d_mint, mint.data(), mint.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));

  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  size_t acc_off = 0;  // doubles
  for (size_t gi = 0; gi < order.size(); ++gi) {
    const GKey &k = order[gi];
    const auto sg = shape_of(k);
    const McShape sh = sg.first;
    const size_t smem = staging(k, sh.blk);
    const int n_units = static_cast<int>(ranges[gi].second);
    HIPCHECK(launch_mc<R>(std::get<0>(k), std::get<1>(k), std::get<2>(k), sh, d_jobs, d_units + ranges[gi].first,
                          n_units, d_counter + 16 * gi, d_accg + acc_off, smem,
                          grid_for(k, sh, sg.second, smem, n_units), c->stream));
    acc_off += static_cast<size_t>(n_units) * 2 * NSTAT * sh.blk;
  }
  {
    int maxw = 0;
    for (int k : nks) maxw = std::max(maxw, 2 * NSTAT * k);
    const int tb = 128;
    dim3 grid((maxw + tb - 1) / tb, static_cast<unsigned>(jobs.size()));
    const int nji = static_cast<int>(nj);
    hipLaunchKernelGGL(reduce_partials_kernel, grid, dim3(tb), 0, c->stream, d_partial, d_sums, d_base, d_mint,
                       d_mint + nj, d_mint + 2 * nj, nji);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  std::vector<double> sums(soff);
  HIPCHECK(hipMemcpyAsync(sums.data(), d_sums, soff * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  device_ms += ms;

  // statistics (MonteCarloPricer.hpp:443-478)
  for (size_t i = 0; i < jobs.size(); ++i) {
    const JobSpec &js = built[i].spec;
    myql_mc_result *res = js.res;
    const int nk = nks[i];
    const double *sm = sums.data() + out_off[i];
    const double invM = 1.0 / static_cast<double>(js.cfg->num_paths);
    const double df = built[i].df;
    auto put = [&](int st, double *val, double *err) {
      for (int k = 0; k < nk; ++k) {
        const double mean = sm[(2 * st) * nk + k] * invM;
        const double var = sm[(2 * st + 1) * nk + k] * invM - mean * mean;
        if (val) val[js.k0 + k] = mean * df;
        if (err) err[js.k0 + k] = std::sqrt(var * invM) * df;
      }
    };
    put(ST_PRICE, res->price, res->price_std_err);
    if (js.mode != MYQL_GREEK_NONE) {
      put(ST_DELTA, res->delta, res->delta_std_err);
      put(ST_GAMMA, res->gamma, res->gamma_std_err);
    }
    if (js.mode == MYQL_GREEK_FULL) {
      put(ST_VEGA0, res->vega[0], res->vega_std_err[0]);
      put(ST_VEGA1, res->vega[1], res->vega_std_err[1]);
      put(ST_THETA, res->theta, res->theta_std_err);
      put(ST_RHO, res->rho, res->rho_std_err);
      put(ST_VANNA0, res->vanna[0], res->vanna_std_err[0]);
      put(ST_VANNA1, res->vanna[1], res->vanna_std_err[1]);
      put(ST_CHARM, res->charm, res->charm_std_err);
    }
    if (js.k0 == 0) res->total_steps = js.cfg->num_paths * js.cfg->time_steps * built[i].variants_alg;
  }
}


// ---------------------------------------------------------------------------
// MLMC: Giles' adaptive driver on the host (MultiLevelMonteCarloPricer.hpp
// :46-241) over level kernels that keep every stream's state on the device.
// ---------------------------------------------------------------------------
// Pinned host memory (hipHostMalloc) of the MLMC graphs' copy nodes.
struct HostBuf {
  void *p = nullptr;
  void reserve(size_t bytes) { HIPCHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  template <typename T>
  T *as() const { return static_cast<T *>(p); }
};

// Weights of the level means 0..L in a multilevel estimate. Giles' MLMC: 1.
// ML2R (multilevel Richardson-Romberg, Lemaire & Pages 2017) on refiners
// n_l = 2^l and weak order alpha: the Richardson weights w_l = L_l(0), the
// Lagrange basis at 0 over the nodes x_l = n_l^-alpha, i.e.
// w_l = prod_{j != l} x_j / (x_j - x_l), which solve sum_l w_l = 1,
// sum_l w_l x_l^k = 0 (k = 1..L); level l's mean E[P_l - P_{l-1}] then
// carries W_l = sum_{j >= l} w_j (W_0 = 1).
std::vector<double> mlmc_level_weights(size_t L, double alpha, bool ml2r) {
  std::vector<double> W(L + 1, 1.0);
  if (!ml2r) return W;
  std::vector<double> x(L + 1), w(L + 1);
  for (size_t l = 0; l <= L; ++l) x[l] = std::pow(2.0, -alpha * static_cast<double>(l));
  for (size_t l = 0; l <= L; ++l) {
    double p = 1.0;
    for (size_t j = 0; j <= L; ++j)
      if (j != l) p *= x[j] / (x[j] - x[l]);
    w[l] = p;
  }
  double t = 0.0;
  for (size_t l = L + 1; l-- > 0;) {
    t += w[l];
    W[l] = t;
  }
  W[0] = 1.0;  // sum of all w, exactly 1 up to rounding
  return W;
}

// One level's launch sequence -- dN upload, level kernel, block-sum download
// -- captured once as a HIP graph and replayed on every visit of the level
// by the Giles loop (the per-visit path count goes through pinned memory).
struct LevelGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  ~LevelGraph() {
    if (exec) (void)hipDeviceSynchronize();  // no replay may be in flight
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
  }
};

template <typename R>
struct MlmcRun {
  myql_ctx *c;
  const myql_model &m;
  const int32_t *schemes;
  const myql_instrument &in;
  int mode;
  const myql_mlmc_config &cfg;
  double S0, r, q;
  uint64_t S;
  double device_ms = 0.0;
  unsigned nb = 0;        // blocks of a level launch
  size_t job_stride = 0;  // bytes per level in c->mlmc_job: Job<R> + strike
  HostBuf h_dn, h_sums;
  DevBuf d_dn, d_sums;
  std::vector<std::unique_ptr<LevelGraph>> graphs;

  void prepare() {
    nb = static_cast<unsigned>((S + BLOCK - 1) / BLOCK);
    const size_t L = cfg.L_max + 1;
    job_stride = (sizeof(Job<R>) + 255) / 256 * 256 + 256;
    c->mlmc_job.reserve(L * job_stride);
    h_dn.reserve(L * sizeof(uint64_t));
    h_sums.reserve(L * MLMC_SUMS * nb * sizeof(double));
    d_dn.reserve(L * sizeof(uint64_t));
    d_sums.reserve(L * MLMC_SUMS * nb * sizeof(double));
    graphs.resize(L);
  }

  // Job of level l (fine/coarse steppers) in its own device slot, and the
  // level's captured launch sequence.
  void build_level(size_t l) {
    Job<R> J;
    fill_instrument(J, m, schemes, in, mode, S0, r, cfg.fd_bump, false);
    const bool pd = in.kind != MYQL_INSTR_EUROPEAN;
    const double T = in.maturity;
    const int nf = m.n_factors;
    const bool coupled = l > 0;
    const uint64_t fs = cfg.base_steps * (1ull << l);
    const uint64_t steps = coupled ? fs / 2 : cfg.base_steps;
    const double dtf = coupled ? T / static_cast<double>(fs) : T / static_cast<double>(cfg.base_steps);
    int nmax[2] = {0, 0};
    for (int f = 0; f < nf; ++f)
      if (schemes[f] == MYQL_SCHEME_NCI) {
        const myql_heston_params &p = m.heston[f];
        nmax[f] = nci_nmax(p, dtf, cfg.flags & MYQL_FLAG_REFERENCE_NCI_ROWS);
        const bool ref = cfg.flags & MYQL_FLAG_REFERENCE_NCI_ROWS;
        NciInfo ni = nci_table(c, 4.0 * p.kappa * p.theta / (p.sigma * p.sigma), nci_nmax(p, dtf, false),
                               cfg.precision, nci_nmax(p, dtf, true));
        set_nci(J, f, ni, ref);
      }
    fill_variant(J.var[0], m, schemes, dtf, r, q, T, coupled ? fs : steps, pd, nmax);
    if (coupled) fill_variant(J.var[1], m, schemes, 2.0 * dtf, r, q, T, steps, pd, nmax);
    J.var_list[0] = 0;
    J.var_list[1] = 1;
    J.n_groups = 1;
    J.group_size[0] = 1;
    J.M = 0;
    J.S = S;
    J.steps = steps;
    J.nk = 1;
    J.nkp = 1;
    char *slot = static_cast<char *>(c->mlmc_job.p) + l * job_stride;
    double *d_k = reinterpret_cast<double *>(slot + (sizeof(Job<R>) + 255) / 256 * 256);
    J.strikes = d_k;
    HIPCHECK(hipMemcpyAsync(slot, &J, sizeof(J), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(d_k, in.strikes, sizeof(double), hipMemcpyHostToDevice, c->stream));
    const size_t smem = static_cast<size_t>(J.n_fields) * MAXV * BLOCK * sizeof(R);
    if (smem > c->lds_max) throw Fail(MYQL_E_UNSUPPORTED, "staging area exceeds the device's LDS per block");
    HIPCHECK(hipStreamSynchronize(c->stream));
    const Job<R> *dj = reinterpret_cast<const Job<R> *>(slot);
    uint64_t *st = c->mlmc_streams.as<uint64_t>();
    double *bs = d_sums.as<double>() + l * MLMC_SUMS * nb;
    const uint64_t *dn = d_dn.as<uint64_t>() + l;
    const int cp = coupled ? 1 : 0;
    auto &G = graphs[l];
    G = std::make_unique<LevelGraph>();
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(d_dn.as<uint64_t>() + l, h_dn.as<uint64_t>() + l, sizeof(uint64_t),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      const bool gk = mode != MYQL_GREEK_NONE;
#define MYQL_MLMC_LAUNCH(NF_, GK_)                                                                              \
  hipLaunchKernelGGL((mlmc_level_kernel<R, NF_, GK_>), dim3(nb), dim3(BLOCK), smem, c->stream, dj, st, dn, cp, bs)
      if (nf == 0) {
        if (gk) MYQL_MLMC_LAUNCH(0, true);
        else MYQL_MLMC_LAUNCH(0, false);
      } else if (nf == 1) {
        if (gk) MYQL_MLMC_LAUNCH(1, true);
        else MYQL_MLMC_LAUNCH(1, false);
      } else {
        if (gk) MYQL_MLMC_LAUNCH(2, true);
        else MYQL_MLMC_LAUNCH(2, false);
      }
#undef MYQL_MLMC_LAUNCH
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_sums.as<double>() + l * MLMC_SUMS * nb, bs, MLMC_SUMS * nb * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream);
    const hipError_t ce = hipStreamEndCapture(c->stream, &G->g);
    HIPCHECK(e);
    HIPCHECK(ce);
    HIPCHECK(hipGraphInstantiate(&G->exec, G->g, nullptr, nullptr, 0));
  }

  // adds the level's new paths into acc (MLMC_SUMS sums: price, delta and
  // gamma differences and their squares)
  void level(size_t l, uint64_t dN, double *acc) {
    if (!graphs[l]) build_level(l);
    h_dn.as<uint64_t>()[l] = dN;
    HIPCHECK(hipEventRecord(c->ev0, c->stream));
    HIPCHECK(hipGraphLaunch(graphs[l]->exec, c->stream));
    HIPCHECK(hipEventRecord(c->ev1, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    device_ms += ms;
    const double *h = h_sums.as<double>() + l * MLMC_SUMS * nb;
    for (unsigned b = 0; b < nb; ++b)
      for (int a = 0; a < MLMC_SUMS; ++a) acc[a] += h[MLMC_SUMS * b + a];
  }

  void run(myql_mlmc_result &res) {
    // persistent per-stream generators (the reference's rng_pool, :55-61)
    const uint64_t *init = stream_states(c, cfg.seed, cfg.stream_layout, S);
    c->mlmc_streams.reserve(S * 32);
    HIPCHECK(hipMemcpyAsync(c->mlmc_streams.p, init, S * 32, hipMemcpyDeviceToDevice, c->stream));
    prepare();
    const size_t Lmax = cfg.L_max;
    size_t L = cfg.L_min;
    const bool ml2r = cfg.estimator == MYQL_MLMC_ML2R;
    std::vector<uint64_t> N(Lmax + 1, 0);
    std::vector<std::array<double, MLMC_SUMS>> acc(Lmax + 1);
    for (auto &a : acc) a.fill(0.0);
    std::vector<double> V(Lmax + 1, 0.0), cost(Lmax + 1, 0.0);
    for (size_t l = 0; l <= Lmax; ++l) cost[l] = static_cast<double>(cfg.base_steps * (1ull << l));
    auto mean = [&](size_t l, int a) { return acc[l][a] / static_cast<double>(N[l]); };
    auto var = [&](size_t l, int a) {
      const double m = mean(l, a);
      return std::max(1e-18, acc[l][a + 1] / static_cast<double>(N[l]) - m * m);
    };
    double alpha = cfg.alpha, beta = cfg.beta;
    const double eps = cfg.epsilon;
    std::vector<double> W = mlmc_level_weights(L, cfg.alpha, ml2r);
    // sample sizes of the weighted sum sum_l W_l mean_l at variance eps^2/2
    // for the least cost (W = 1: Giles' N_l, MultiLevelMonteCarloPricer.hpp:
    // 88-95)
    auto optimal = [&](size_t l, size_t Lc) -> uint64_t {
      if (!(V[l] > 1e-18)) return 0;
      double sumVC = 0.0;
      for (size_t k = 0; k <= Lc; ++k) sumVC += std::fabs(W[k]) * std::sqrt(V[k] * cost[k]);
      return static_cast<uint64_t>(
          std::ceil(2.0 * std::pow(eps, -2.0) * std::fabs(W[l]) * std::sqrt(V[l] / cost[l]) * sumVC));
    };
    bool converged = false;
    while (!converged) {
      W = mlmc_level_weights(L, cfg.alpha, ml2r);
      for (size_t l = 0; l <= L; ++l) {
        const uint64_t opt = optimal(l, L);
        const uint64_t pilot = std::max<uint64_t>(20, cfg.N0 / (1ull << l));
        const uint64_t target = std::max(pilot, opt);
        if (N[l] < target) {
          const uint64_t dN = target - N[l];
          level(l, dN, acc[l].data());
          N[l] += dN;
          if (N[l] > 1) V[l] = var(l, 0);
        }
      }
      if (cfg.adaptive_rates && L >= 2) {
        auto order = [&](const std::vector<double> &vals) {
          double sl = 0, sy = 0, sly = 0, sl2 = 0;
          int cnt = 0;
          const size_t l0 = L - 2;
          for (size_t l = l0; l <= L; ++l)
            if (vals[l] > 1e-16) {
              const double y = std::log(vals[l]), x = static_cast<double>(l - l0);
              sl += x;
              sy += y;
              sly += x * y;
              sl2 += x * x;
              ++cnt;
            }
          if (cnt < 2) return -1.0;
          return -((cnt * sly - sl * sy) / (cnt * sl2 - sl * sl)) / std::log(2.0);
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
        // Giles' bias estimate (MultiLevelMonteCarloPricer.hpp:190-205)
        double eff = std::fabs(mean(L, 0));
        if (L > 0) eff = std::max(eff, 0.5 * std::fabs(mean(L - 1, 0)));
        bias = eff / (std::pow(2.0, alpha) - 1.0);
      } else if (L == 0) {
        bias = __builtin_inf();  // one depth: nothing to compare
      } else {
        // the change from depth L - 1 to depth L, which the deeper
        // extrapolation removes: an estimate of the shallower one's bias,
        // so an upper estimate of this one's
        const std::vector<double> Wp = mlmc_level_weights(L - 1, cfg.alpha, true);
        double d = W[L] * mean(L, 0);
        for (size_t l = 0; l < L; ++l) d += (W[l] - Wp[l]) * mean(l, 0);
        bias = std::fabs(d);
      }
      if (bias > eps / 1.4142135623730951 && L < Lmax) ++L;
      else converged = true;
    }
    W = mlmc_level_weights(L, cfg.alpha, ml2r);
    const double df = std::exp(-r * in.maturity);
    double est[3] = {0.0, 0.0, 0.0}, ev[3] = {0.0, 0.0, 0.0};
    res.total_steps = 0;
    res.alpha_estimated = alpha;
    res.beta_estimated = beta;
    res.n_levels = static_cast<int32_t>(L + 1);
    const bool gk = mode != MYQL_GREEK_NONE;
    for (size_t l = 0; l <= L; ++l) {
      const uint64_t spp = (l == 0) ? cfg.base_steps : (cfg.base_steps * (1ull << l) * 3 / 2);
      res.total_steps += N[l] * spp;
      if (res.paths_per_level) res.paths_per_level[l] = N[l];
      if (res.mean_diffs) res.mean_diffs[l] = mean(l, 0) * df;
      if (res.var_diffs) res.var_diffs[l] = V[l] * df * df;
      if (res.level_weights) res.level_weights[l] = W[l];
      for (int a = 0; a < (gk ? 3 : 1); ++a) {
        est[a] += W[l] * mean(l, 2 * a);
        ev[a] += W[l] * W[l] * (a == 0 ? V[l] : var(l, 2 * a)) / static_cast<double>(N[l]);
      }
    }
    res.price = est[0] * df;
    res.price_std_err = std::sqrt(ev[0]) * df;
    res.delta = est[1] * df;
    res.delta_std_err = std::sqrt(ev[1]) * df;
    res.gamma = est[2] * df;
    res.gamma_std_err = std::sqrt(ev[2]) * df;
    res.device_ms = device_ms;
  }
};

void mlmc_validate(const myql_model *m, const int32_t *schemes, const myql_instrument *in, const myql_mlmc_config *cfg) {
  myql_mc_config probe{};
  probe.num_paths = 1;
  probe.time_steps = 1;
  probe.precision = cfg ? cfg->precision : MYQL_FP64;
  validate(m, schemes, in, &probe);
  if (cfg->L_max > 30 || cfg->L_min > cfg->L_max) throw Fail(MYQL_E_INVALID_ARG, "need L_min <= L_max <= 30");
  if (cfg->base_steps < 1) throw Fail(MYQL_E_INVALID_ARG, "base_steps must be >= 1");
  if (!(cfg->epsilon > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "epsilon must be positive");
  if (cfg->estimator != MYQL_MLMC_GILES && cfg->estimator != MYQL_MLMC_ML2R)
    throw Fail(MYQL_E_INVALID_ARG, "estimator must be MYQL_MLMC_GILES or MYQL_MLMC_ML2R");
  if (cfg->estimator == MYQL_MLMC_ML2R && !(cfg->alpha > 0.0 && cfg->alpha <= 8.0))
    throw Fail(MYQL_E_INVALID_ARG, "ML2R needs the weak order alpha in (0, 8]");
}

// ---------------------------------------------------------------------------
// learned control variate (cv_kernel.hpp)
// ---------------------------------------------------------------------------
void cv_validate(const myql_mc_job *jobs, int64_t n, const myql_cv_config *cfg) {
  if (!jobs || n < 1 || !cfg) throw Fail(MYQL_E_INVALID_ARG, "control variate: no jobs or config");
  if (n > CV_MAXJOBS) throw Fail(MYQL_E_UNSUPPORTED, "control variate: at most 16 jobs per call");
  if (cfg->points < 1 || cfg->points > MYQL_CV_MAX_POINTS) throw Fail(MYQL_E_INVALID_ARG, "points must be in 1..16");
  const myql_mc_job &a = jobs[0];
  for (int64_t i = 0; i < n; ++i) {
    const myql_mc_job &b = jobs[i];
    if (!b.model || !b.instrument || !b.config || !b.result) throw Fail(MYQL_E_INVALID_ARG, "null job field");
    validate(b.model, b.schemes, b.instrument, b.config);
    if (std::memcmp(b.model, a.model, sizeof(myql_model)) != 0 || b.schemes[0] != a.schemes[0] ||
        b.schemes[1] != a.schemes[1])
      throw Fail(MYQL_E_INVALID_ARG, "control variate: all jobs must share model and schemes");
    const myql_instrument &ib = *b.instrument, &ia = *a.instrument;
    if (ib.kind != MYQL_INSTR_EUROPEAN) throw Fail(MYQL_E_UNSUPPORTED, "control variate: European options only");
    if (ib.payoff != ia.payoff || ib.option_type != ia.option_type)
      throw Fail(MYQL_E_INVALID_ARG, "control variate: all jobs must share payoff and option type");
    if (ib.n_strikes > CV_OUT) throw Fail(MYQL_E_UNSUPPORTED, "control variate: at most 64 strikes per job");
    if (ib.n_strikes != ia.n_strikes ||
        std::memcmp(ib.strikes, ia.strikes, static_cast<size_t>(ib.n_strikes) * sizeof(double)) != 0)
      throw Fail(MYQL_E_INVALID_ARG, "control variate: all jobs must share the strike strip (one head output per strike)");
    if (b.S0 != a.S0 || b.r != a.r || b.q != a.q) throw Fail(MYQL_E_INVALID_ARG, "control variate: jobs must share S0, r, q");
    if (b.greek_mode == MYQL_GREEK_FULL)
      throw Fail(MYQL_E_UNSUPPORTED, "control variate: GreekMode::Full is not supported (None or Essential)");
    if (b.greek_mode != a.greek_mode) throw Fail(MYQL_E_INVALID_ARG, "control variate: jobs must share the Greek mode");
    if (b.config->flags & MYQL_FLAG_REFERENCE_NCI_ROWS)
      throw Fail(MYQL_E_UNSUPPORTED, "control variate: the reference NCI table semantics are not supported");
  }
}

CvNorm cv_norm(const myql_mc_job *jobs, int64_t n) {
  double Tmax = 0.0;
  for (int64_t i = 0; i < n; ++i) Tmax = std::max(Tmax, jobs[i].instrument->maturity);
  const myql_model &m = *jobs[0].model;
  const double vbar = std::max(expected_avg_variance(m, Tmax), 1e-6);
  CvNorm nm;
  nm.ts = static_cast<float>(1.0 / Tmax);
  nm.xs = static_cast<float>(1.0 / std::sqrt(vbar * Tmax));
  nm.vs0 = m.n_factors > 0 ? static_cast<float>(1.0 / m.heston[0].theta) : 0.0f;
  nm.vs1 = m.n_factors > 1 ? static_cast<float>(1.0 / m.heston[1].theta) : 0.0f;
  nm.logS0 = static_cast<float>(std::log(jobs[0].S0));
  nm.S0 = static_cast<float>(jobs[0].S0);
  return nm;
}

// Owner of one recorded option set.
struct CvRecord {
  CvSet set{};
  std::vector<unsigned char> job;  // host copy of the Job<R> the paths run
  DevBuf rec, st, sm, strikes, mu, isd;
};

// The set's terminal-jump expectation E[P_k | S_T-] (vanilla, sharp and
// smoothed digital payoffs under Merton or Kou jumps; S.payoff, S.call,
// S.smooth, S.eps_scale, S.strikes and S.nk set; ks the strikes on the
// host): its parameters and, on the log-spot grid around S0, its table
// (cached in the context, built on the context stream when new).
// The terminal-jump table of set S (parameters filled), from the context's
// cache or built on the context stream into a new cache entry.
const double *cv_rb_table(myql_ctx *c, const myql_model &m, double T, double S0, const double *ks,
                          const CvSet &S) {
  std::vector<double> key{static_cast<double>(m.jump_kind), m.jump[0], m.jump[1], m.jump[2], m.jump[3], T, S0,
                          static_cast<double>(S.payoff), static_cast<double>(S.call),
                          static_cast<double>(S.smooth), S.smooth ? S.eps_scale : 0.0};
  key.insert(key.end(), ks, ks + S.nk);
  auto it = c->cv_rb.find(key);
  if (it != c->cv_rb.end()) return it->second->as<double>();
  auto buf = std::make_unique<DevBuf>();
  buf->reserve(static_cast<size_t>(S.nk) * CV_RB_PTS * 2 * sizeof(double));
  HIPCHECK(launch_cv_rb_table(S, buf->as<double>(), c->stream));
  const double *p = buf->as<double>();
  c->cv_rb.emplace(std::move(key), std::move(buf));
  return p;
}

// 16-point Gauss-Legendre rule on [-1, 1]: roots of P_16 by Newton's
// method from the Chebyshev-like first guesses, w = 2 / ((1 - x^2) P_16'(x)^2)
static void gauss_legendre16(double *x, double *w) {
  constexpr int n = 16;
  auto legendre = [](double z, double &dp) {
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; ++j) {
      const double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    dp = n * (z * p1 - p2) / (z * z - 1.0);
    return p1;
  };
  for (int i = 0; i < n; ++i) {
    double z = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), dp;
    for (int it = 0; it < 100; ++it) {
      const double dz = legendre(z, dp) / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    legendre(z, dp);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

void cv_terminal_jump(myql_ctx *c, const myql_model &m, double T, double S0, const double *ks, CvSet &S) {
  const int nk = S.nk;
  gauss_legendre16(S.gl_x, S.gl_w);
  // Merton terminal jump: E[P | S_T-] in closed form for vanilla and sharp
  // digital payoffs: J = m0 + n mu + sqrt(n) delta Z, n ~ Poisson(lambda T),
  // m0 the drift compensation (JumpPolicies.hpp:55-135)
  // Kou terminal jump (JumpPolicies.hpp:55-135): up and down jump counts are
  // independent Poisson(lambda T p), Poisson(lambda T (1 - p)); sums of
  // exponentials, so E[P | S_T-] has a closed form at integer shapes
  S.rb_on = (m.jump_kind == MYQL_JUMP_MERTON || m.jump_kind == MYQL_JUMP_KOU) ? 1 : 0;
  // smoothed digitals: the transition window [K - 10 eps, K + 10 eps] must
  // lie at positive spots
  if (S.smooth) {
    const double eps = 0.70710678118654752440 / S.eps_scale;
    for (int k = 0; k < nk; ++k)
      if (!(ks[k] - 10.0 * eps > 0.0) || !std::isfinite(eps)) S.rb_on = 0;
  }
  S.rb_kou = 0;
  S.rb_nd = 0;
  if (S.rb_on && m.jump_kind == MYQL_JUMP_KOU) {
    const double lam = m.jump[0], p = m.jump[1], e1 = m.jump[2], e2 = m.jump[3];
    const double kap = (p * e1) / (e1 - 1.0) + ((1.0 - p) * e2) / (e2 + 1.0) - 1.0;
    S.rb_kou = 1;
    S.kou_e1 = e1;
    S.kou_e2 = e2;
    S.rb_m0 = -lam * kap * T;
    S.rb_p0 = std::exp(-lam * T);
    // Poisson weights of one side until their mass is covered
    auto weights = [](double mean, double *w, int &n) {
      double pn = std::exp(-mean), mass = 0.0;
      n = 0;
      while (n < CV_RB_TERMS && mass < 1.0 - 1e-17) {
        w[n] = pn;
        mass += pn;
        ++n;
        pn *= mean / n;
      }
      return mass;
    };
    double wu[CV_RB_TERMS], wd[CV_RB_TERMS];
    const double mu = weights(lam * T * p, wu, S.rb_n), md = weights(lam * T * (1.0 - p), wd, S.rb_nd);
    double su = 0.0, sd = 0.0, cu = 1.0, cd = 1.0;
    for (int j = 0; j < S.rb_n; ++j, cu *= e1 / (e1 - 1.0)) su += wu[j] * cu;
    for (int j = 0; j < S.rb_nd; ++j, cd *= e2 / (e2 + 1.0)) sd += wd[j] * cd;
    // series coefficients G_r = sum_l NB_l W_{l+r} of P(A - B > t) =
    // sum_r G_r Q_r(eA t), W_j = wA_j cA^j, NB_l = sum_m wB_m cB^m
    // C(m-1+l, l) a^l b^m, a = eA / (eA + eB), b = 1 - a; per orientation
    // (A the up or the down side) and tilt
    auto series = [](const double *wA, int nA, double cA, const double *wB, int nB, double cB, double eA,
                     double eB, double *G) {
      const double a = eA / (eA + eB), b = eB / (eA + eB);
      double W[CV_RB_TERMS], NB[CV_RB_TERMS];
      double cj = 1.0;
      for (int j = 0; j < nA; ++j, cj *= cA) {
        W[j] = wA[j] * cj;
        NB[j] = 0.0;
      }
      double vm = 1.0;
      for (int m = 0; m < nB; ++m, vm *= b * cB) {
        double nb = wB[m] * vm;  // V_m b^m C(m - 1, 0)
        for (int l = 0; l < nA; ++l) {
          NB[l] += nb;
          nb *= (static_cast<double>(m + l) / static_cast<double>(l + 1)) * a;
        }
      }
      for (int r = 0; r < CV_RB_TERMS; ++r) G[r] = 0.0;
      for (int r = 1; r < nA; ++r)
        for (int l = 0; l + r < nA; ++l) G[r] += NB[l] * W[l + r];
    };
    series(wu, S.rb_n, 1.0, wd, S.rb_nd, 1.0, e1, e2, S.kou_g[0]);
    series(wu, S.rb_n, e1 / (e1 - 1.0), wd, S.rb_nd, e2 / (e2 + 1.0), e1 - 1.0, e2 + 1.0, S.kou_g[1]);
    series(wd, S.rb_nd, 1.0, wu, S.rb_n, 1.0, e2, e1, S.kou_g[2]);
    series(wd, S.rb_nd, e2 / (e2 + 1.0), wu, S.rb_n, e1 / (e1 - 1.0), e2 + 1.0, e1 - 1.0, S.kou_g[3]);
    S.kou_mass = mu * md - S.rb_p0;
    S.kou_m1 = su * sd - S.rb_p0;
    // the n >= 1 sums stay unbiased only where both sides' mass is covered
    // and e^J has a mean (e1 > 1)
    const double mass = mu * md;
    if (!(e1 > 1.0) || mass < 1.0 - 1e-14) S.rb_on = 0;
    else {
      S.rb_y0 = std::log(S0) - 3.0;
      S.rb_h = 5.0 / (CV_RB_PTS - 1);
      S.rb_inv_h = 1.0 / S.rb_h;
      S.rb = cv_rb_table(c, m, T, S0, ks, S);
    }
  } else if (S.rb_on) {
    const double lt = m.jump[0] * T, mu = m.jump[1], dl = m.jump[2];
    S.rb_m0 = -m.jump[0] * (std::exp(mu + 0.5 * dl * dl) - 1.0) * T;
    S.rb_p0 = std::exp(-lt);
    double pn = S.rb_p0, mass = S.rb_p0;
    S.rb_n = 0;
    for (int n = 1; n <= CV_RB_TERMS && mass < 1.0 - 1e-17; ++n) {
      pn *= lt / n;
      mass += pn;
      S.rb_p[S.rb_n] = pn;
      S.rb_m[S.rb_n] = S.rb_m0 + n * mu;
      S.rb_s[S.rb_n] = std::sqrt(static_cast<double>(n)) * dl;
      ++S.rb_n;
    }
    S.kou_mass = 0.0;
    S.kou_m1 = 0.0;
    for (int n = 0; n < S.rb_n; ++n) {
      S.kou_mass += S.rb_p[n];
      S.kou_m1 += S.rb_p[n] * std::exp(S.rb_m[n] - S.rb_m0 + 0.5 * S.rb_s[n] * S.rb_s[n]);
    }
    // counts beyond the tabulated terms would bias E[P | S_T-]: then the
    // estimator goes without that term (D = 0), and stays unbiased
    if (mass < 1.0 - 1e-14) {
      S.rb_on = 0;
    } else {
      S.rb_y0 = std::log(S0) - 3.0;
      S.rb_h = 5.0 / (CV_RB_PTS - 1);
      S.rb_inv_h = 1.0 / S.rb_h;
      S.rb = cv_rb_table(c, m, T, S0, ks, S);
    }
  }
}

// Set up the recording of the base paths of job `jb` under config `cfg`
// (paths, seed, stream layout, precision) at `points` control points: the
// job, the buffers and the set's terminal-jump expectation tables. The paths
// themselves are recorded by cv_record_paths, batched over sets.
template <typename R>
void cv_record(myql_ctx *c, const myql_mc_job &jb, const myql_mc_config &cfg, int points, CvRecord &out) {
  const myql_model &m = *jb.model;
  const myql_instrument &in = *jb.instrument;
  const int nk = static_cast<int>(in.n_strikes);
  JobSpec js{jb.model, jb.schemes, jb.instrument, jb.greek_mode, &cfg, jb.S0, jb.r, jb.q, nullptr, 0, nk};
  Job<R> J{};
  std::vector<double> ks;
  build_job<R>(c, js, J, ks);
  J.strikes = nullptr;
  J.partial = nullptr;
  out.job.assign(reinterpret_cast<const unsigned char *>(&J), reinterpret_cast<const unsigned char *>(&J + 1));
  const uint64_t M = cfg.num_paths, steps = cfg.time_steps;
  const double T = in.maturity, dt = T / static_cast<double>(steps);
  CvSet &S = out.set;
  S = CvSet{};
  S.M = M;
  S.S = J.S;
  S.Ms = (M + J.S - 1) / J.S * J.S;
  const int nf = m.n_factors;
  S.nj = nf == 0 ? 1 : static_cast<int>(std::min<uint64_t>(points, steps));
  for (int j = 0; j <= S.nj; ++j) S.step_of[j] = static_cast<int>((static_cast<uint64_t>(j) * steps) / S.nj);
  for (int j = 0; j < S.nj; ++j) {
    const double tj = S.step_of[j] * dt;
    S.t_of[j] = static_cast<float>(tj);
    S.disc_of[j] = std::exp(jb.r * (T - tj));
  }
  S.nf = nf;
  for (int f = 0; f < nf; ++f) {
    const myql_heston_params &p = m.heston[f];
    S.c2x[f] = static_cast<float>((p.kappa * p.rho / p.sigma - 0.5) * dt * 0.5 + p.rho / p.sigma);
    S.nci_mean[f] = jb.schemes[f] == MYQL_SCHEME_NCI ? J.nci_mean[f] : nullptr;
    S.nci_rows[f] = jb.schemes[f] == MYQL_SCHEME_NCI ? J.nci_rows[f] : 0;
  }
  out.strikes.reserve(nk * sizeof(double));
  HIPCHECK(hipMemcpyAsync(out.strikes.p, in.strikes, nk * sizeof(double), hipMemcpyHostToDevice, c->stream));
  S.strikes = out.strikes.as<double>();
  S.nk = nk;
  S.payoff = in.payoff;
  S.call = in.option_type == MYQL_CALL ? 1 : 0;
  S.mode = jb.greek_mode;
  S.smooth = (jb.greek_mode != MYQL_GREEK_NONE && in.payoff != MYQL_PAYOFF_VANILLA) ? 1 : 0;
  S.eps_scale = J.eps_scale;
  S.mult_up = J.mult_up;
  S.mult_dn = J.mult_dn;
  S.inv2h = J.inv2h;
  S.invh2 = J.invh2;
  S.T = T;
  S.df = std::exp(-jb.r * T);
  out.rec.reserve(static_cast<size_t>(S.nj) * CV_REC * S.Ms * sizeof(float));
  out.st.reserve(S.Ms * sizeof(double));
  S.rec = out.rec.as<float>();
  S.ST = out.st.as<double>();
  S.Sm = nullptr;
  if (m.jump_kind != MYQL_JUMP_NONE) {
    out.sm.reserve(S.Ms * sizeof(double));
    S.Sm = out.sm.as<double>();
  }
  cv_terminal_jump(c, m, T, jb.S0, in.strikes, S);
}

// Device copies of sets (kept alive by the caller until it synchronises).
const CvSet *cv_upload_sets(myql_ctx *c, const std::vector<CvSet> &sets, DevBuf &d_sets) {
  d_sets.reserve(sets.size() * sizeof(CvSet));
  HIPCHECK(hipMemcpyAsync(d_sets.p, sets.data(), sets.size() * sizeof(CvSet), hipMemcpyHostToDevice, c->stream));
  return d_sets.as<CvSet>();
}

// Order of the sets for batched grids: by factor count (one grid each), the
// most work per thread first, so the short sets fill the tail.
std::vector<size_t> cv_batch_order(const std::vector<CvSet> &sets) {
  std::vector<size_t> order(sets.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  auto work = [&](size_t i) {
    const CvSet &t = sets[i];
    return static_cast<double>(t.step_of[t.nj]) * static_cast<double>(t.Ms / t.S);
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (sets[a].nf != sets[b].nf) return sets[a].nf < sets[b].nf;
    return work(a) > work(b);
  });
  return order;
}

// Record the base paths of recs (all of precision R), one grid per factor
// count.
template <typename R>
void cv_record_paths(myql_ctx *c, const std::vector<CvRecord *> &recs) {
  DevBuf &d_jobs = c->cv_jobs, &d_sets = c->cv_sets, &d_tasks = c->cv_tasks;
  const size_t n = recs.size();
  if (n == 0) return;
  std::vector<CvSet> sets(n);
  std::vector<Job<R>> hj(n);
  for (size_t i = 0; i < n; ++i) {
    sets[i] = recs[i]->set;
    std::memcpy(static_cast<void *>(&hj[i]), recs[i]->job.data(), sizeof(Job<R>));
  }
  const CvSet *ds = cv_upload_sets(c, sets, d_sets);
  d_jobs.reserve(n * sizeof(Job<R>));
  HIPCHECK(hipMemcpyAsync(d_jobs.p, hj.data(), n * sizeof(Job<R>), hipMemcpyHostToDevice, c->stream));
  const std::vector<size_t> order = cv_batch_order(sets);
  std::vector<CvPathTask> tasks(n);
  std::vector<std::tuple<int, size_t, size_t, unsigned>> grids;  // nf, first task, tasks, blocks
  for (size_t g0 = 0; g0 < n;) {
    const int nf = sets[order[g0]].nf;
    size_t g1 = g0;
    unsigned blocks = 0;
    for (; g1 < n && sets[order[g1]].nf == nf; ++g1) {
      const size_t i = order[g1];
      tasks[g1] = CvPathTask{static_cast<int>(i), static_cast<int>(i), blocks};
      blocks += static_cast<unsigned>((sets[i].S + BLOCK - 1) / BLOCK);
    }
    grids.emplace_back(nf, g0, g1 - g0, blocks);
    g0 = g1;
  }
  d_tasks.reserve(n * sizeof(CvPathTask));
  HIPCHECK(hipMemcpyAsync(d_tasks.p, tasks.data(), n * sizeof(CvPathTask), hipMemcpyHostToDevice, c->stream));
  for (const auto &g : grids)
    HIPCHECK(launch_cv_paths<R>(std::get<0>(g), d_jobs.as<Job<R>>(), ds, d_tasks.as<CvPathTask>() + std::get<1>(g),
                                static_cast<int>(std::get<2>(g)), std::get<3>(g), c->stream));
}

int cv_eval_blocks(myql_ctx *c, uint64_t M);

// Evaluate sets (device copies ds) in `mode` (0 price, 1 price + plain
// delta/gamma, 2 fitting sums), coefficients coefs[i] ([CV_NREG][64] on the
// host; null in mode 2), one grid per factor count; returns the sums, nk x ns
// per set at sums[i * 64 * ns], once the device has finished.
void cv_eval_sets(myql_ctx *c, int mode, const DevBuf &prm, const CvNorm &nm, const std::vector<CvSet> &sets,
                  const CvSet *ds, const std::vector<const double *> &coefs, std::vector<double> &sums) {
  const size_t n = sets.size();
  const int ns = mode == 2 ? CV_NS_FIT : (mode == 1 ? CV_NS_PRICE : 4);
  const std::vector<size_t> order = cv_batch_order(sets);
  std::vector<int> blocks(n);
  size_t part_n = 0;
  for (size_t i = 0; i < n; ++i) {
    blocks[i] = cv_eval_blocks(c, sets[i].M);
    part_n += static_cast<size_t>(blocks[i]) * (BLOCK / 64) * CV_OUT * ns;
  }
  DevBuf &part = c->cv_part, &d_sums = c->cv_sums, &d_coef = c->cv_coef, &d_tasks = c->cv_tasks;
  part.reserve(part_n * sizeof(double));
  d_sums.reserve(n * CV_OUT * ns * sizeof(double));
  if (mode != 2) {
    std::vector<double> h(n * CV_NREG * CV_OUT);
    for (size_t i = 0; i < n; ++i) std::memcpy(&h[i * CV_NREG * CV_OUT], coefs[i], CV_NREG * CV_OUT * sizeof(double));
    d_coef.reserve(h.size() * sizeof(double));
    HIPCHECK(hipMemcpyAsync(d_coef.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  std::vector<CvEvalTask> tasks(n);
  std::vector<std::tuple<int, size_t, size_t, unsigned, int>> grids;  // nf, first, count, blocks, max nk
  size_t off = 0;
  for (size_t g0 = 0; g0 < n;) {
    const int nf = sets[order[g0]].nf;
    size_t g1 = g0;
    unsigned first = 0;
    int max_nk = 0;
    for (; g1 < n && sets[order[g1]].nf == nf; ++g1) {
      const size_t i = order[g1];
      tasks[g1] = CvEvalTask{static_cast<int>(i), first, static_cast<unsigned>(blocks[i] * (BLOCK / 64)),
                             mode == 2 ? nullptr : d_coef.as<double>() + i * CV_NREG * CV_OUT, part.as<double>() + off,
                             d_sums.as<double>() + i * CV_OUT * ns};
      off += static_cast<size_t>(blocks[i]) * (BLOCK / 64) * CV_OUT * ns;
      first += static_cast<unsigned>(blocks[i]);
      max_nk = std::max(max_nk, sets[i].nk);
    }
    grids.emplace_back(nf, g0, g1 - g0, first, max_nk);
    g0 = g1;
  }
  d_tasks.reserve(n * sizeof(CvEvalTask));
  HIPCHECK(hipMemcpyAsync(d_tasks.p, tasks.data(), n * sizeof(CvEvalTask), hipMemcpyHostToDevice, c->stream));
  for (const auto &g : grids)
    HIPCHECK(launch_cv_eval(std::get<0>(g), mode, prm.as<float>(), nm, ds, d_tasks.as<CvEvalTask>() + std::get<1>(g),
                            static_cast<int>(std::get<2>(g)), std::get<3>(g), std::get<4>(g), c->stream));
  sums.assign(n * CV_OUT * ns, 0.0);
  HIPCHECK(hipMemcpyAsync(sums.data(), d_sums.p, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

void cv_init_params(uint32_t seed, std::vector<float> &p) {
  p.assign(CV_NPARAM, 0.0f);
  uint64_t x = 0x5DEECE66Dull ^ seed;
  auto u = [&]() {  // splitmix64 -> [0, 1)
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1p-53;
  };
  auto fill = [&](int off, int fan_in, int fan_out, int count) {  // Glorot uniform
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < count; ++i) p[off + i] = static_cast<float>((2.0 * u() - 1.0) * a);
  };
  fill(CV_W1, CV_IN, CV_H, CV_H * CV_IN);
  fill(CV_W2, CV_H, CV_H, CV_H * CV_H);
  fill(CV_W3, CV_H, CV_H, CV_H * CV_H);
  fill(CV_W4, CV_H, CV_OUT, CV_OUT * CV_H);
}

int cv_eval_blocks(myql_ctx *c, uint64_t M) {
  return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(static_cast<uint64_t>(c->n_cu) * 2,
                                                                    (M + 127) / 128)));
}

// Per-strike least-squares coefficients of P on the regressors (M_A, M_V,
// M_J, D) from the fitting sums of products of (1, P, R0..R3); regressors
// without variance are left out.
void cv_fit(const std::vector<double> &sums, int nk, uint64_t M, double *coef /* [CV_NREG][64] */) {
  const double invM = 1.0 / static_cast<double>(M);
  for (int k = 0; k < nk; ++k) {
    double S[6][6];
    int q = 0;
    for (int u = 0; u < 6; ++u)
      for (int w = u; w < 6; ++w) S[u][w] = S[w][u] = sums[static_cast<size_t>(k) * CV_NS_FIT + q++] * invM;
    // covariances: index 1 = P, 2..5 = regressors (index 0 is the constant)
    double C[4][4], b[4];
    int act[4], na = 0;
    for (int i = 0; i < 4; ++i) {
      const double var = S[2 + i][2 + i] - S[0][2 + i] * S[0][2 + i];
      if (var > 1e-300 && var > 1e-24 * (S[1][1] + 1e-300)) act[na++] = i;
    }
    for (int a = 0; a < na; ++a) {
      const int i = act[a];
      b[a] = S[1][2 + i] - S[0][1] * S[0][2 + i];
      for (int e = 0; e < na; ++e) {
        const int j = act[e];
        C[a][e] = S[2 + i][2 + j] - S[0][2 + i] * S[0][2 + j];
      }
      C[a][a] *= 1.0 + 1e-12;
    }
    // Gaussian elimination with partial pivoting
    double x[4] = {0, 0, 0, 0};
    for (int col = 0; col < na; ++col) {
      int piv = col;
      for (int rw = col + 1; rw < na; ++rw)
        if (std::fabs(C[rw][col]) > std::fabs(C[piv][col])) piv = rw;
      for (int e = 0; e < na; ++e) std::swap(C[col][e], C[piv][e]);
      std::swap(b[col], b[piv]);
      if (!(std::fabs(C[col][col]) > 0.0)) continue;
      for (int rw = col + 1; rw < na; ++rw) {
        const double f = C[rw][col] / C[col][col];
        for (int e = col; e < na; ++e) C[rw][e] -= f * C[col][e];
        b[rw] -= f * b[col];
      }
    }
    for (int rw = na - 1; rw >= 0; --rw) {
      double s = b[rw];
      for (int e = rw + 1; e < na; ++e) s -= C[rw][e] * x[e];
      x[rw] = std::fabs(C[rw][rw]) > 0.0 ? s / C[rw][rw] : 0.0;
    }
    for (int i = 0; i < CV_NREG; ++i) coef[i * CV_OUT + k] = 0.0;
    for (int a = 0; a < na; ++a) coef[act[a] * CV_OUT + k] = std::isfinite(x[a]) ? x[a] : 0.0;
  }
}

void cv_train(myql_ctx *c, const myql_mc_job *jobs, int64_t n, const myql_cv_config *cfg, myql_cv_net *net) {
  if (cfg->batch < CV_TROWS || cfg->batch % CV_TROWS != 0) throw Fail(MYQL_E_INVALID_ARG, "batch must be a multiple of 32");
  if (cfg->steps < 1 || cfg->train_paths < 1) throw Fail(MYQL_E_INVALID_ARG, "need steps >= 1 and train_paths >= 1");
  const CvNorm nm = cv_norm(jobs, n);
  const int nk = static_cast<int>(jobs[0].instrument->n_strikes);
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  while (c->cv_recs.size() < static_cast<size_t>(n)) c->cv_recs.push_back(std::make_unique<CvRecord>());
  std::vector<CvSet> sets;
  const int mblocks = 64;
  DevBuf mpart;
  mpart.reserve(static_cast<size_t>(nk) * mblocks * 2 * sizeof(double));
  std::vector<CvRecord *> rp;
  for (int64_t i = 0; i < n; ++i) {
    myql_mc_config tc = *jobs[i].config;
    tc.num_paths = cfg->train_paths;
    tc.seed = cfg->train_seed;
    tc.precision = MYQL_FP32;
    tc.num_streams = 0;
    cv_record<float>(c, jobs[i], tc, static_cast<int>(cfg->points), *c->cv_recs[i]);
    rp.push_back(c->cv_recs[i].get());
  }
  cv_record_paths<float>(c, rp);
  for (int64_t i = 0; i < n; ++i) {
    CvRecord &rc = *c->cv_recs[i];
    // per-strike scale of the training targets
    HIPCHECK(launch_cv_moments(rc.set, mpart.as<double>(), mblocks, c->stream));
    std::vector<double> h(static_cast<size_t>(nk) * mblocks * 2);
    HIPCHECK(hipMemcpyAsync(h.data(), mpart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    std::vector<float> mu(nk), isd(nk);
    const double invM = 1.0 / static_cast<double>(rc.set.M);
    for (int k = 0; k < nk; ++k) {
      double s = 0.0, s2 = 0.0;
      for (int b = 0; b < mblocks; ++b) {
        s += h[2 * (static_cast<size_t>(k) * mblocks + b)];
        s2 += h[2 * (static_cast<size_t>(k) * mblocks + b) + 1];
      }
      const double mean = s * invM, var = std::max(s2 * invM - mean * mean, 0.0);
      mu[k] = static_cast<float>(mean);
      isd[k] = var > 1e-20 ? static_cast<float>(1.0 / std::sqrt(var)) : 0.0f;
    }
    rc.mu.reserve(nk * sizeof(float));
    rc.isd.reserve(nk * sizeof(float));
    HIPCHECK(hipMemcpyAsync(rc.mu.p, mu.data(), nk * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(rc.isd.p, isd.data(), nk * sizeof(float), hipMemcpyHostToDevice, c->stream));
    rc.set.mu = rc.mu.as<float>();
    rc.set.inv_sd = rc.isd.as<float>();
    sets.push_back(rc.set);
  }
  DevBuf d_sets, prm, mom, var, gpart, lpart, hist, step;
  d_sets.reserve(sets.size() * sizeof(CvSet));
  HIPCHECK(hipMemcpyAsync(d_sets.p, sets.data(), sets.size() * sizeof(CvSet), hipMemcpyHostToDevice, c->stream));
  std::vector<float> p0;
  cv_init_params(cfg->init_seed, p0);
  const int nb = static_cast<int>(cfg->batch / CV_TROWS);
  prm.reserve(CV_NPARAM * sizeof(float));
  mom.reserve(CV_NPARAM * sizeof(float));
  var.reserve(CV_NPARAM * sizeof(float));
  gpart.reserve(static_cast<size_t>(nb) * CV_NPARAM * sizeof(float));
  lpart.reserve(nb * sizeof(float));
  hist.reserve(cfg->steps * sizeof(float));
  step.reserve(sizeof(int));
  HIPCHECK(hipMemcpyAsync(prm.p, p0.data(), CV_NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemsetAsync(mom.p, 0, CV_NPARAM * sizeof(float), c->stream));
  HIPCHECK(hipMemsetAsync(var.p, 0, CV_NPARAM * sizeof(float), c->stream));
  HIPCHECK(hipMemsetAsync(step.p, 0, sizeof(int), c->stream));
  CvTrainArgs a{prm.as<float>(), mom.as<float>(), var.as<float>(), gpart.as<float>(), lpart.as<float>(),
                hist.as<float>(), d_sets.as<CvSet>(), static_cast<int>(n), nm, cfg->train_seed, step.as<int>(),
                static_cast<int>(cfg->batch), nk, static_cast<float>(cfg->lr),
                static_cast<float>(1.0 / (static_cast<double>(cfg->batch) * nk)), true, nullptr, nullptr};
  // one Adam step (train, reduce + update, counter) captured as a graph and
  // replayed; the step counter lives on the device, so replays need no
  // argument changes
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t le = launch_cv_train_step(a, c->stream);
  const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  HIPCHECK(le);
  HIPCHECK(ce);
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(graph);
    HIPCHECK(ie);
  }
  hipError_t ge = hipSuccess;
  for (uint32_t it = 0; it < cfg->steps && ge == hipSuccess; ++it) ge = hipGraphLaunch(exec, c->stream);
  // the replays must have finished before their executable graph is
  // destroyed: the runtime (and a profiler's tracing thread) still reads a
  // queued launch's dispatch data from it
  const hipError_t se = hipStreamSynchronize(c->stream);
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  HIPCHECK(ge);
  HIPCHECK(se);
  // per-(maturity, strike) coefficients of the regressors, fitted on the
  // training paths (independent of any pricing paths)
  std::memset(net->coef, 0, sizeof(net->coef));
  {
    std::vector<double> fs;
    cv_eval_sets(c, 2, prm, nm, sets, d_sets.as<CvSet>(), std::vector<const double *>(sets.size(), nullptr), fs);
    for (int64_t i = 0; i < n; ++i) {
      const std::vector<double> h(fs.begin() + i * CV_OUT * CV_NS_FIT,
                                  fs.begin() + i * CV_OUT * CV_NS_FIT + static_cast<size_t>(nk) * CV_NS_FIT);
      cv_fit(h, nk, sets[i].M, &net->coef[i][0][0]);
      net->maturity[i] = jobs[i].instrument->maturity;
    }
  }
  net->n_jobs = static_cast<int32_t>(n);
  net->n_strikes = nk;
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  std::vector<float> h(cfg->steps);
  HIPCHECK(hipMemcpyAsync(net->params, prm.p, CV_NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(h.data(), hist.p, cfg->steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  const uint32_t tail = std::max<uint32_t>(1, std::min<uint32_t>(50, cfg->steps));
  double l = 0.0;
  for (uint32_t i = cfg->steps - tail; i < cfg->steps; ++i) l += h[i];
  net->final_loss = l / tail;
  net->train_ms = ms;
  const float nv[6] = {nm.ts, nm.xs, nm.vs0, nm.vs1, nm.logS0, nm.S0};
  std::memcpy(net->norm, nv, sizeof(nv));
}

// Per-strike results of a priced set from its sums (ns = 4 or 8 per strike).
void cv_results(const myql_mc_job &jb, const CvSet &S, int mode, const double *sums, double device_ms,
                myql_cv_result *cr) {
  const int ns = mode == 1 ? CV_NS_PRICE : 4;
  const uint64_t M = S.M;
  const double invM = 1.0 / static_cast<double>(M), df = S.df;
  myql_mc_result &res = *jb.result;
  for (int k = 0; k < S.nk; ++k) {
    const double *s = sums + static_cast<size_t>(k) * ns;
    const double mc = s[0] * invM, vc = std::max(s[1] * invM - mc * mc, 0.0);
    const double mp = s[2] * invM, vp = std::max(s[3] * invM - mp * mp, 0.0);
    const bool use = vc < vp;
    if (res.price) res.price[k] = (use ? mc : mp) * df;
    if (res.price_std_err) res.price_std_err[k] = std::sqrt((use ? vc : vp) * invM) * df;
    if (mode == 1) {
      const double md = s[4] * invM, vd = s[5] * invM - md * md, mg = s[6] * invM, vg = s[7] * invM - mg * mg;
      if (res.delta) res.delta[k] = md * df;
      if (res.delta_std_err) res.delta_std_err[k] = std::sqrt(vd * invM) * df;
      if (res.gamma) res.gamma[k] = mg * df;
      if (res.gamma_std_err) res.gamma_std_err[k] = std::sqrt(vg * invM) * df;
    }
    if (cr) {
      if (cr->plain_price) cr->plain_price[k] = mp * df;
      if (cr->plain_std_err) cr->plain_std_err[k] = std::sqrt(vp * invM) * df;
      if (cr->var_ratio) cr->var_ratio[k] = vc > 0.0 ? vp / vc : (vp > 0.0 ? INFINITY : 1.0);
      if (cr->used_cv) cr->used_cv[k] = use ? 1 : 0;
    }
  }
  res.total_steps = M * jb.config->time_steps;
  res.device_ms = device_ms;
}

// Device bytes of a job's recording.
size_t cv_record_bytes(const myql_mc_job &jb, int points) {
  const uint64_t M = jb.config->num_paths;
  const uint64_t nj = jb.model->n_factors == 0 ? 1 : std::min<uint64_t>(points, jb.config->time_steps);
  return static_cast<size_t>(M) * (nj * CV_REC * sizeof(float) + 2 * sizeof(double));
}

// The jobs are priced in chunks whose recordings together fit kCvChunkBytes
// of device memory (at least one job per chunk): per chunk, every job's
// paths in one grid per precision and factor count, every set's evaluation
// in one grid per factor count and Greek mode, one synchronisation. Each
// job's device_ms is the whole call's device time.
constexpr size_t kCvChunkBytes = size_t(32) << 30;

void cv_price(myql_ctx *c, const myql_mc_job *jobs, int64_t n, const myql_cv_config *cfg, const myql_cv_net *net,
              myql_cv_result *cr) {
  if (net->n_strikes != jobs[0].instrument->n_strikes || net->n_jobs < 1 || net->n_jobs > CV_MAXJOBS)
    throw Fail(MYQL_E_INVALID_ARG, "control variate: the network was trained for another strike strip");
  std::vector<const double *> coef(n);
  for (int64_t i = 0; i < n; ++i) {
    // the coefficients fitted for this maturity
    int t = -1;
    for (int j = 0; j < net->n_jobs; ++j)
      if (net->maturity[j] == jobs[i].instrument->maturity) t = j;
    if (t < 0) throw Fail(MYQL_E_INVALID_ARG, "control variate: no coefficients trained for this maturity");
    coef[i] = &net->coef[t][0][0];
  }
  CvNorm nm;
  nm.ts = net->norm[0];
  nm.xs = net->norm[1];
  nm.vs0 = net->norm[2];
  nm.vs1 = net->norm[3];
  nm.logS0 = net->norm[4];
  nm.S0 = net->norm[5];
  const int points = static_cast<int>(cfg->points);
  DevBuf prm;
  prm.reserve(CV_NPARAM * sizeof(float));
  HIPCHECK(hipMemcpyAsync(prm.p, net->params, CV_NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  // per job: its set and sums, kept until the device time is known
  std::vector<CvSet> all_sets(n);
  std::vector<std::vector<double>> all_sums(n);
  std::vector<int> modes(n);
  for (int64_t i0 = 0; i0 < n;) {
    int64_t i1 = i0;
    size_t bytes = 0;
    while (i1 < n && (i1 == i0 || bytes + cv_record_bytes(jobs[i1], points) <= kCvChunkBytes))
      bytes += cv_record_bytes(jobs[i1++], points);
    while (c->cv_recs.size() < static_cast<size_t>(i1 - i0)) c->cv_recs.push_back(std::make_unique<CvRecord>());
    std::vector<CvRecord *> r32, r64;
    for (int64_t i = i0; i < i1; ++i) {
      CvRecord &rc = *c->cv_recs[i - i0];
      if (jobs[i].config->precision == MYQL_FP32) {
        cv_record<float>(c, jobs[i], *jobs[i].config, points, rc);
        r32.push_back(&rc);
      } else {
        cv_record<double>(c, jobs[i], *jobs[i].config, points, rc);
        r64.push_back(&rc);
      }
      modes[i] = jobs[i].greek_mode == MYQL_GREEK_ESSENTIAL ? 1 : 0;
    }
    cv_record_paths<float>(c, r32);
    cv_record_paths<double>(c, r64);
    for (int mode = 0; mode < 2; ++mode) {
      std::vector<CvSet> sets;
      std::vector<const double *> cf;
      std::vector<int64_t> idx;
      for (int64_t i = i0; i < i1; ++i)
        if (modes[i] == mode) {
          sets.push_back(c->cv_recs[i - i0]->set);
          cf.push_back(coef[i]);
          idx.push_back(i);
        }
      if (sets.empty()) continue;
      const CvSet *ds = cv_upload_sets(c, sets, c->cv_sets);
      std::vector<double> sums;
      cv_eval_sets(c, mode, prm, nm, sets, ds, cf, sums);
      const int ns = mode == 1 ? CV_NS_PRICE : 4;
      for (size_t q = 0; q < idx.size(); ++q) {
        all_sets[idx[q]] = sets[q];
        all_sums[idx[q]].assign(sums.begin() + q * CV_OUT * ns, sums.begin() + (q + 1) * CV_OUT * ns);
      }
    }
    i0 = i1;
  }
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  for (int64_t i = 0; i < n; ++i)
    cv_results(jobs[i], all_sets[i], modes[i], all_sums[i].data(), ms, cr ? cr + i : nullptr);
}

int fail(myql_ctx *c, const Fail &e) {
  if (c) c->err = e.what();
  return e.code;
}

// Split jobs into strike chunks of at most 64 and run them by precision.
// A quasi-Monte Carlo call's replicates: each prices num_paths / R points
// under its own digital shift into its own result arrays; the caller's
// result gets their means and standard errors.
struct QmcRep {
  myql_mc_config cfg;
  myql_mc_result res;
  std::vector<std::vector<double>> store;
};

// value / standard-error field pairs of a result
using ResField = double *myql_mc_result::*;
const std::pair<ResField, ResField> kResPairs[] = {
    {&myql_mc_result::price, &myql_mc_result::price_std_err}, {&myql_mc_result::delta, &myql_mc_result::delta_std_err},
    {&myql_mc_result::gamma, &myql_mc_result::gamma_std_err}, {&myql_mc_result::theta, &myql_mc_result::theta_std_err},
    {&myql_mc_result::rho, &myql_mc_result::rho_std_err},     {&myql_mc_result::charm, &myql_mc_result::charm_std_err}};

void run_specs(myql_ctx *c, const std::vector<JobSpec> &in_specs) {
  std::vector<JobSpec> s32, s64;
  std::vector<std::vector<std::unique_ptr<QmcRep>>> reps(in_specs.size());
  auto push = [&](const JobSpec &s) {
    const int nk = static_cast<int>(s.instr->n_strikes);
    for (int k0 = 0; k0 < nk; k0 += STRIKE_CHUNK) {
      JobSpec t = s;
      t.k0 = k0;
      t.nk = std::min(STRIKE_CHUNK, nk - k0);
      (s.cfg->precision == MYQL_FP32 ? s32 : s64).push_back(t);
    }
  };
  for (size_t si = 0; si < in_specs.size(); ++si) {
    const JobSpec &s = in_specs[si];
    validate(s.model, s.schemes, s.instr, s.cfg);
    if (!s.res) throw Fail(MYQL_E_INVALID_ARG, "null result");
    s.res->device_ms = 0.0;
    if (s.cfg->stream_layout != MYQL_STREAMS_SOBOL) {
      push(s);
      continue;
    }
    const uint64_t R = qmc_reps(*s.cfg);
    const size_t nk = static_cast<size_t>(s.instr->n_strikes);
    for (uint64_t r = 0; r < R; ++r) {
      auto q = std::make_unique<QmcRep>();
      q->cfg = *s.cfg;
      q->cfg.num_paths = s.cfg->num_paths / R;
      q->res = myql_mc_result{};
      auto give = [&](double *const &user, double *&mine) {
        if (!user) return;
        q->store.emplace_back(nk, 0.0);
        mine = q->store.back().data();
      };
      for (const auto &pr : kResPairs) {
        give(s.res->*pr.first, q->res.*pr.first);
        give(s.res->*pr.second, q->res.*pr.second);
      }
      for (int f = 0; f < 2; ++f) {
        give(s.res->vega[f], q->res.vega[f]);
        give(s.res->vega_std_err[f], q->res.vega_std_err[f]);
        give(s.res->vanna[f], q->res.vanna[f]);
        give(s.res->vanna_std_err[f], q->res.vanna_std_err[f]);
      }
      JobSpec t = s;
      t.cfg = &q->cfg;
      t.res = &q->res;
      t.qmc_rep = static_cast<int>(r);
      reps[si].push_back(std::move(q));
      push(t);
    }
  }
  double ms32 = 0.0, ms64 = 0.0;
  if (!s32.empty()) run_jobs<float>(c, s32, ms32);
  if (!s64.empty()) run_jobs<double>(c, s64, ms64);
  for (size_t si = 0; si < in_specs.size(); ++si) {
    const JobSpec &s = in_specs[si];
    s.res->device_ms = ms32 + ms64;
    if (reps[si].empty()) continue;
    // replicate means, and their standard deviation / sqrt(R) as the error
    const auto &rs = reps[si];
    const double R = static_cast<double>(rs.size());
    const size_t nk = static_cast<size_t>(s.instr->n_strikes);
    auto combine_arr = [&](double *val, double *err, auto get) {
      if (!val) return;
      for (size_t k = 0; k < nk; ++k) {
        double m = 0.0;
        for (const auto &q : rs) m += get(*q)[k];
        m /= R;
        double v = 0.0;
        for (const auto &q : rs) {
          const double d = get(*q)[k] - m;
          v += d * d;
        }
        val[k] = m;
        if (err) err[k] = rs.size() > 1 ? std::sqrt(v / (R - 1.0) / R) : 0.0;
      }
    };
    // the fields the Greek mode produces (the others are left alone, as in
    // a pseudo-random call): price; delta, gamma; the Full Greeks
    const int n_pairs = s.mode == MYQL_GREEK_FULL ? 6 : (s.mode == MYQL_GREEK_ESSENTIAL ? 3 : 1);
    for (int p = 0; p < n_pairs; ++p) {
      const auto &pr = kResPairs[p];
      combine_arr(s.res->*pr.first, s.res->*pr.second, [&](const QmcRep &q) { return q.res.*pr.first; });
    }
    if (s.mode == MYQL_GREEK_FULL)
      for (int f = 0; f < 2; ++f) {
        combine_arr(s.res->vega[f], s.res->vega_std_err[f], [&](const QmcRep &q) { return q.res.vega[f]; });
        combine_arr(s.res->vanna[f], s.res->vanna_std_err[f], [&](const QmcRep &q) { return q.res.vanna[f]; });
      }
    uint64_t steps = 0;
    for (const auto &q : rs) steps += q->res.total_steps;
    s.res->total_steps = steps;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Gil-Pelaez Fourier pricer: options on the device, results assembled on the
// host as FourierPricer::price_scalar does (FourierPricer.hpp:105-404).
// ---------------------------------------------------------------------------
namespace {

void fourier_validate(const myql_fourier_job &j) {
  if (!j.model || !j.instrument || !j.result) throw Fail(MYQL_E_INVALID_ARG, "null argument");
  const myql_model &m = *j.model;
  const myql_instrument &in = *j.instrument;
  if (m.n_factors < 0 || m.n_factors > 2) throw Fail(MYQL_E_INVALID_ARG, "n_factors must be 0, 1 or 2");
  if (m.jump_kind < MYQL_JUMP_NONE || m.jump_kind > MYQL_JUMP_KOU) throw Fail(MYQL_E_INVALID_ARG, "bad jump kind");
  if (m.jump_kind == MYQL_JUMP_KOU && !(m.jump[2] > 1.0)) throw Fail(MYQL_E_INVALID_ARG, "Kou eta1 must exceed 1");
  for (int f = 0; f < m.n_factors; ++f)
    if (!(m.heston[f].sigma > 0.0) || !(m.heston[f].v0 >= 0.0)) throw Fail(MYQL_E_INVALID_ARG, "bad Heston factor");
  if (in.kind != MYQL_INSTR_EUROPEAN) throw Fail(MYQL_E_UNSUPPORTED, "the Fourier pricer prices European options");
  if (in.payoff < MYQL_PAYOFF_VANILLA || in.payoff > MYQL_PAYOFF_ASSET)
    throw Fail(MYQL_E_INVALID_ARG, "bad payoff");
  if (!(in.maturity > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "maturity must be positive");
  if (in.n_strikes <= 0 || !in.strikes) throw Fail(MYQL_E_INVALID_ARG, "no strikes");
  for (int64_t k = 0; k < in.n_strikes; ++k)
    if (!(in.strikes[k] > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "strikes must be positive");
  if (!(j.S0 > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "S0 must be positive");
  if (j.greek_mode < MYQL_GREEK_NONE || j.greek_mode > MYQL_GREEK_FULL) throw Fail(MYQL_E_INVALID_ARG, "bad greek mode");
}

void fourier_run(myql_ctx *c, const myql_fourier_job *jobs, int64_t n_jobs) {
  std::vector<FourierOption> opts;
  for (int64_t ji = 0; ji < n_jobs; ++ji) {
    const myql_fourier_job &j = jobs[ji];
    fourier_validate(j);
    const myql_model &m = *j.model;
    const myql_instrument &in = *j.instrument;
    const double start = j.config ? j.config->start_bound_guess : 20.0;
    FourierOption o{};
    o.model.nf = m.n_factors;
    o.model.jump_kind = m.jump_kind;
    o.model.vol = m.vol;
    for (int f = 0; f < 2; ++f) {
      const myql_heston_params &h = m.heston[f];
      const double hp[5] = {h.kappa, h.theta, h.sigma, h.rho, h.v0};
      for (int e = 0; e < 5; ++e) o.model.h[f][e] = hp[e];
    }
    for (int e = 0; e < 4; ++e) o.model.jump[e] = m.jump[e];
    o.T = in.maturity;
    o.drift = (j.r - j.q) * in.maturity;
    o.start_bound = start / (in.maturity < 1.0 ? std::sqrt(in.maturity) : 1.0);
    o.mode = j.greek_mode;
    for (int64_t k = 0; k < in.n_strikes; ++k) {
      o.logK = std::log(in.strikes[k] / j.S0);
      opts.push_back(o);
    }
  }
  const int n = static_cast<int>(opts.size());
  const int ni = fourier_n_integrals();
  c->fopts.reserve(std::max<size_t>(1, opts.size()) * sizeof(FourierOption));
  c->fout.reserve(std::max<size_t>(1, opts.size()) * ni * sizeof(double));
  HIPCHECK(hipMemcpyAsync(c->fopts.p, opts.data(), opts.size() * sizeof(FourierOption), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  HIPCHECK(launch_fourier(c->fopts.as<FourierOption>(), n, c->fout.as<double>(), c->stream));
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  std::vector<double> I(opts.size() * ni);
  HIPCHECK(hipMemcpyAsync(I.data(), c->fout.p, I.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));

  // integral layout (fourier_launch.hpp): [P1 | P2] x {P, Dx, Dxx, V0, V1, T, VDx0, VDx1, TDx}
  enum { P = 0, DX, DXX, V0, V1, TH, VDX0, VDX1, TDX, NI };
  size_t oi = 0;
  for (int64_t ji = 0; ji < n_jobs; ++ji) {
    const myql_fourier_job &j = jobs[ji];
    const myql_model &m = *j.model;
    const myql_instrument &in = *j.instrument;
    myql_mc_result &res = *j.result;
    const double S0 = j.S0, r = j.r, q = j.q, T = in.maturity;
    const double df_r = std::exp(-r * T), df_q = std::exp(-q * T);
    const bool call = in.option_type == MYQL_CALL;
    const double sign = call ? 1.0 : -1.0;
    const int nf = m.n_factors;
    double chain[2] = {1.0, 1.0};  // vega per unit of sqrt(v0) (AffineTraits vega_chain_factor)
    for (int f = 0; f < nf; ++f) chain[f] = 2.0 * std::sqrt(m.heston[f].v0);
    const bool ess = j.greek_mode != MYQL_GREEK_NONE, full = j.greek_mode == MYQL_GREEK_FULL;
    for (int64_t k = 0; k < in.n_strikes; ++k, ++oi) {
      const double K = in.strikes[k];
      const double *I1 = I.data() + oi * ni, *I2 = I1 + NI;
      const double P1 = 0.5 + I1[P], P2 = 0.5 + I2[P];
      auto put = [&](double *arr, double v) {
        if (arr) arr[k] = v;
      };
      double price = 0, delta = 0, gamma = 0, vega[2] = {0, 0}, theta = 0, rho = 0, vanna[2] = {0, 0}, charm = 0;
      if (in.payoff == MYQL_PAYOFF_VANILLA) {
        price = call ? std::max(0.0, S0 * df_q * P1 - K * df_r * P2)
                     : std::max(0.0, K * df_r * (1.0 - P2) - S0 * df_q * (1.0 - P1));
        delta = call ? df_q * P1 : df_q * (P1 - 1.0);
        gamma = df_q * I1[DX] / S0;
        for (int f = 0; f < std::max(nf, 1); ++f) {
          const double d1 = I1[V0 + f], d2 = I2[V0 + f];
          vega[f] = (S0 * df_q * d1 - K * df_r * d2) * chain[f];
          vanna[f] = df_q * d1 * chain[f];
        }
        const double dP1dT = I1[TH], dP2dT = I2[TH];
        // dP/dT. For puts the reference writes K df_r (r (1 - P2) - dP2/dT)
        // (FourierPricer.hpp:183-184); the T-derivative of K e^{-rT} (1 - P2)
        // carries -r, which its own Monte Carlo theta estimator agrees with,
        // so the discount term takes that sign here.
        theta = call ? S0 * df_q * (-q * P1 + dP1dT) + K * df_r * (r * P2 - dP2dT)
                     : K * df_r * (-r * (1.0 - P2) - dP2dT) + S0 * df_q * (q * (1.0 - P1) + dP1dT);
        const double dP1dr = T * I1[DX], dP2dr = T * I2[DX];
        const double disc = call ? T * K * df_r * P2 : -T * K * df_r * (1.0 - P2);
        rho = S0 * df_q * dP1dr + disc - K * df_r * dP2dr;
        charm = call ? df_q * (-q * P1 + dP1dT) : df_q * (-q * (P1 - 1.0) + dP1dT);
      } else if (in.payoff == MYQL_PAYOFF_CASH) {
        const double base = call ? P2 : 1.0 - P2;
        price = df_r * base;
        delta = sign * df_r * I2[DX] / S0;
        gamma = sign * df_r * (I2[DXX] - I2[DX]) / (S0 * S0);
        for (int f = 0; f < std::max(nf, 1); ++f) {
          vega[f] = sign * df_r * I2[V0 + f] * chain[f];
          vanna[f] = sign * df_r * I2[VDX0 + f] / S0 * chain[f];
        }
        theta = df_r * (sign * I2[TH] - r * base);
        rho = df_r * (sign * T * I2[DX] - T * base);
        charm = sign * df_r * (I2[TDX] - r * I2[DX]) / S0;
      } else {  // asset-or-nothing
        const double base = call ? P1 : 1.0 - P1;
        price = S0 * df_q * base;
        delta = df_q * (base + sign * I1[DX]);
        gamma = sign * df_q * (I1[DX] + I1[DXX]) / S0;
        for (int f = 0; f < std::max(nf, 1); ++f) {
          vega[f] = S0 * df_q * sign * I1[V0 + f] * chain[f];
          vanna[f] = df_q * sign * (I1[V0 + f] + I1[VDX0 + f]) * chain[f];
        }
        theta = S0 * df_q * (-q * base + sign * I1[TH]);
        rho = S0 * df_q * sign * T * I1[DX];
        charm = df_q * (-q * (base + sign * I1[DX]) + sign * I1[TH] + sign * I1[TDX]);
      }
      put(res.price, price);
      if (ess) {
        put(res.delta, delta);
        put(res.gamma, gamma);
      }
      if (full) {
        put(res.vega[0], vega[0]);
        put(res.vega[1], nf >= 2 ? vega[1] : 0.0);
        put(res.theta, theta);
        put(res.rho, rho);
        put(res.vanna[0], vanna[0]);
        put(res.vanna[1], nf >= 2 ? vanna[1] : 0.0);
        put(res.charm, charm);
      }
    }
    res.device_ms = ms;
  }
}

}  // namespace

namespace myql {
bool mc_supported(int nf, int nv, bool pd) {
  // lockstep widths: Europeans simulate base, vol bumps and maturity bumps
  // (5, or 7 with two factors) and derive the rate bumps; path-dependent
  // lockstep adds the two rate-bumped paths (7 or 9)
  if (nv == 1) return nf >= 0 && nf <= 2;
  if (!pd) return nv == (nf == 2 ? 7 : 5);
  return nv == (nf == 2 ? 9 : 7);
}
}  // namespace myql

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

myql_ctx *myql_create(int device) {
  auto *c = new (std::nothrow) myql_ctx();
  if (!c) return nullptr;
  try {
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Fail(MYQL_E_INVALID_ARG, "no such device");
    c->device = device;
    set_device(c);
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    // static LDS of the kernels (queue slot, valid flags, reduction) is < 4 KiB
    c->lds_max = prop.sharedMemPerBlock > 4096 ? prop.sharedMemPerBlock - 4096 : prop.sharedMemPerBlock;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->ev0));
    HIPCHECK(hipEventCreate(&c->ev1));
    {
      // this device's copy of the no-allocation scratch array: its address
      // is looked up once per device, with the device current, and kept
      static std::mutex mu;
      static std::map<int, void *> addr;
      std::lock_guard<std::mutex> lock(mu);
      auto it = addr.find(device);
      if (it == addr.end()) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        it = addr.emplace(device, p).first;
      }
      c->scratch = it->second;
    }
    c->jump_polys = host_jump_polys(64);
    c->d_polys.reserve(c->jump_polys.size() * 8);
    HIPCHECK(hipMemcpy(c->d_polys.p, c->jump_polys.data(), c->jump_polys.size() * 8, hipMemcpyHostToDevice));
    return c;
  } catch (const Fail &) {
    delete c;
    return nullptr;
  }
}

void myql_destroy(myql_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  c->streams.clear();
  c->nci.clear();
  c->cv_rb.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *myql_last_error(const myql_ctx *c) { return c ? c->err.c_str() : "null context"; }

int myql_tuning_report(const myql_ctx *c, char *buf, int64_t len) {
  if (!c || !buf || len <= 0) return MYQL_E_INVALID_ARG;
  std::string out;
  for (const auto &kv : c->tuned) {
    const int prec = std::get<0>(kv.first), nf = std::get<1>(kv.first), nv = std::get<2>(kv.first);
    const bool pd = std::get<3>(kv.first);
    char line[160];
    std::snprintf(line, sizeof line,
                  "%s nf=%d nv=%d %s: block %d unroll %d grid %dx%s; trials (block/unroll/grid[/w4: four-wave "
                  "variant]) ",
                  prec == MYQL_FP32 ? "fp32" : "fp64", nf, nv, pd ? "path-dependent" : "european",
                  kv.second.shape.blk, kv.second.shape.unr, kv.second.gridf,
                  kv.second.shape.occ ? " four-wave" : "");
    out += line;
    out += kv.second.trials;
    out += "\n";
  }
  const size_t n = std::min(out.size(), static_cast<size_t>(len - 1));
  std::memcpy(buf, out.data(), n);
  buf[n] = '\0';
  return MYQL_OK;
}

int myql_mc_calculate(myql_ctx *c, const myql_model *model, const int32_t schemes[2], const myql_instrument *instr,
                      int32_t greek_mode, const myql_mc_config *cfg, double S0, double r, double q,
                      myql_mc_result *res) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (greek_mode < MYQL_GREEK_NONE || greek_mode > MYQL_GREEK_FULL) throw Fail(MYQL_E_INVALID_ARG, "bad greek mode");
    JobSpec s{model, schemes, instr, greek_mode, cfg, S0, r, q, res, 0, 0};
    run_specs(c, {s});
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_mc_calculate_noalloc(myql_ctx *c, const myql_model *model, const int32_t schemes[2],
                              const myql_instrument *instr, int32_t greek_mode, const myql_mc_config *cfg, double S0,
                              double r, double q, myql_mc_result *res) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    // no cache trimming here (that frees): only make the device current
    HIPCHECK(hipSetDevice(c->device));
    if (!c->scratch) throw Fail(MYQL_E_HIP, "myql_mc_calculate_noalloc: no scratch address for this device");
    if (greek_mode < MYQL_GREEK_NONE || greek_mode > MYQL_GREEK_FULL) throw Fail(MYQL_E_INVALID_ARG, "bad greek mode");
    if (cfg && cfg->stream_layout == MYQL_STREAMS_SOBOL)
      throw Fail(MYQL_E_UNSUPPORTED, "myql_mc_calculate_noalloc: quasi-Monte Carlo tables are allocated (use "
                                     "myql_mc_calculate)");
    struct Mode {
      myql_ctx *c;
      explicit Mode(myql_ctx *x) : c(x) {
        c->noalloc = true;
        c->scratch_off = 0;
      }
      ~Mode() { c->noalloc = false; }
    } mode(c);
    JobSpec s{model, schemes, instr, greek_mode, cfg, S0, r, q, res, 0, 0};
    run_specs(c, {s});
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_mc_calculate_batch(myql_ctx *c, const myql_mc_job *jobs, int64_t n_jobs) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) throw Fail(MYQL_E_INVALID_ARG, "bad job list");
    std::vector<JobSpec> specs;
    for (int64_t i = 0; i < n_jobs; ++i) {
      const myql_mc_job &j = jobs[i];
      if (j.greek_mode < MYQL_GREEK_NONE || j.greek_mode > MYQL_GREEK_FULL) throw Fail(MYQL_E_INVALID_ARG, "bad greek mode");
      specs.push_back({j.model, j.schemes, j.instrument, j.greek_mode, j.config, j.S0, j.r, j.q, j.result, 0, 0});
    }
    if (!specs.empty()) run_specs(c, specs);
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_fourier_calculate_batch(myql_ctx *c, const myql_fourier_job *jobs, int64_t n_jobs) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) throw Fail(MYQL_E_INVALID_ARG, "bad job list");
    if (n_jobs > 0) fourier_run(c, jobs, n_jobs);
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_mlmc_calculate(myql_ctx *c, const myql_model *model, const int32_t schemes[2], const myql_instrument *instr,
                        int32_t greek_mode, const myql_mlmc_config *cfg, double S0, double r, double q,
                        myql_mlmc_result *res) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!model || !schemes || !instr || !cfg || !res) throw Fail(MYQL_E_INVALID_ARG, "null argument");
    mlmc_validate(model, schemes, instr, cfg);
    const uint64_t S = cfg->num_streams ? cfg->num_streams : 65536;
    if (cfg->precision == MYQL_FP32) {
      MlmcRun<float> run{c, *model, schemes, *instr, greek_mode, *cfg, S0, r, q, S};
      run.run(*res);
    } else {
      MlmcRun<double> run{c, *model, schemes, *instr, greek_mode, *cfg, S0, r, q, S};
      run.run(*res);
    }
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_stream_draws(myql_ctx *c, uint64_t seed, int32_t layout, uint64_t n_streams, uint64_t n_per,
                            uint64_t *out) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!out || n_streams == 0) throw Fail(MYQL_E_INVALID_ARG, "bad arguments");
    const uint64_t *st = stream_states(c, seed, layout, n_streams);
    DevBuf o;
    o.reserve(n_streams * n_per * 8 + 8);
    const unsigned nb = static_cast<unsigned>((n_streams + 255) / 256);
    hipLaunchKernelGGL(debug_draws_kernel, dim3(nb), dim3(256), 0, c->stream, st, n_streams, n_per, o.as<uint64_t>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, o.p, n_streams * n_per * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

static int debug_map(myql_ctx *c, int which, const double *a, const double *b, int64_t n, double *o64, float *o32) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (n <= 0) return MYQL_OK;
    DevBuf da, db, d64, d32;
    da.reserve(n * 8);
    db.reserve(n * 8);
    d64.reserve(n * 8);
    d32.reserve(n * 4);
    HIPCHECK(hipMemcpyAsync(da.p, a, n * 8, hipMemcpyHostToDevice, c->stream));
    if (b) HIPCHECK(hipMemcpyAsync(db.p, b, n * 8, hipMemcpyHostToDevice, c->stream));
    const unsigned nb = static_cast<unsigned>((n + 255) / 256);
    if (which == 0)
      hipLaunchKernelGGL(debug_normcdfinv_kernel, dim3(nb), dim3(256), 0, c->stream, da.as<double>(), n,
                         d64.as<double>(), d32.as<float>());
    else if (which == 1)
      hipLaunchKernelGGL(debug_poissinv_kernel, dim3(nb), dim3(256), 0, c->stream, da.as<double>(), db.as<double>(), n,
                         d64.as<double>());
    else
      hipLaunchKernelGGL(debug_chi2_kernel, dim3(nb), dim3(256), 0, c->stream, da.as<double>(), db.as<double>(), n,
                         d64.as<double>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(o64, d64.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (o32) HIPCHECK(hipMemcpyAsync(o32, d32.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

// The device trackers and calculate_to_buffer on given log-price paths: one
// lane per path runs trk_init, trk_update after every step and
// finalize_stage as the FP64 path kernel does, then eval3 per strike.
constexpr int DBG_PATH_BLK = 64;
__global__ void __launch_bounds__(DBG_PATH_BLK)
    debug_path_payoffs_kernel(const Job<double> *__restrict__ J, const double *__restrict__ logS, int n_paths,
                              int n_steps, double *__restrict__ base, double *__restrict__ up,
                              double *__restrict__ dn) {
  extern __shared__ __align__(16) unsigned char smem[];
  double *stage = reinterpret_cast<double *>(smem);
  const int tid = threadIdx.x;
  const int path = blockIdx.x * DBG_PATH_BLK + tid;
  if (path >= n_paths) return;  // each lane stages and reads only its own slots
  const VarParams<double> *P = &J->var[V_BASE];
  const double dt = P->dt;
  const double *x = logS + static_cast<size_t>(path) * (n_steps + 1);
  Trk<double> tr;
  trk_init(J, tr, x[0], dt);
  for (int k = 1; k <= n_steps; ++k) trk_update(J, tr, x[k], dt);
  finalize_stage<double, DBG_PATH_BLK>(J, tr, x[n_steps], dt, P->total_time, stage, V_BASE, tid);
  const bool bumps = J->mode != 0;
  for (int j = 0; j < J->nk; ++j) {
    double p = 0.0, pu = 0.0, pd = 0.0;
    eval3<double, DBG_PATH_BLK>(J, stage, V_BASE, tid, J->strikes[j], bumps, p, pu, pd);
    const size_t o = static_cast<size_t>(path) * J->nk + j;
    base[o] = p;
    up[o] = pu;
    dn[o] = pd;
  }
}

int myql_debug_path_payoffs(myql_ctx *c, const myql_instrument *in, int32_t greek_mode, double S0, double fd_bump,
                            const double *logS, int64_t n_paths, int64_t n_steps, double *base, double *up,
                            double *dn) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!in || !logS || !base || !up || !dn || n_paths <= 0 || n_steps <= 0 || n_paths > (1 << 20) ||
        n_steps > (1 << 20))
      throw Fail(MYQL_E_INVALID_ARG, "debug_path_payoffs: bad arguments");
    // the paths' own dt is maturity / n_steps, as in a pricing call; a
    // zero-factor model only carries the instrument through build_job
    myql_model m{};
    m.n_factors = 0;
    m.vol = 0.2;
    m.jump_kind = MYQL_JUMP_NONE;
    const int32_t schemes[2] = {MYQL_SCHEME_NULL, MYQL_SCHEME_NULL};
    myql_mc_config cfg{};
    cfg.num_paths = static_cast<uint64_t>(n_paths);
    cfg.time_steps = static_cast<uint64_t>(n_steps);
    cfg.seed = 1;
    cfg.fd_bump = fd_bump;
    cfg.vol_bump = cfg.t_bump = cfg.r_bump = 1e-4;
    cfg.precision = MYQL_FP64;
    cfg.stream_layout = MYQL_STREAMS_JUMP;
    cfg.num_streams = 1;
    validate(&m, schemes, in, &cfg);
    // no model here, so no automatic bandwidth: a smoothed digital needs
    // its bandwidth given
    if (greek_mode != MYQL_GREEK_NONE && in->payoff != MYQL_PAYOFF_VANILLA && !(in->payoff_eps > 0.0))
      throw Fail(MYQL_E_INVALID_ARG, "debug_path_payoffs: a digital payoff in a Greek mode needs payoff_eps > 0");
    JobSpec js{&m, schemes, in, greek_mode, &cfg, S0, 0.0, 0.0, nullptr, 0, static_cast<int>(in->n_strikes)};
    Job<double> J;
    std::vector<double> ks;
    build_job<double>(c, js, J, ks);
    const size_t nk = ks.size(), nx = static_cast<size_t>(n_paths) * (n_steps + 1),
                 no = static_cast<size_t>(n_paths) * nk;
    DevBuf dj, dk, dx, dout;
    dj.reserve(sizeof(Job<double>));
    dk.reserve(nk * 8);
    dx.reserve(nx * 8);
    dout.reserve(3 * no * 8);
    J.strikes = dk.as<double>();
    HIPCHECK(hipMemcpyAsync(dk.p, ks.data(), nk * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(dj.p, &J, sizeof(J), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(dx.p, logS, nx * 8, hipMemcpyHostToDevice, c->stream));
    const size_t smem = static_cast<size_t>(J.n_fields) * MAXV * DBG_PATH_BLK * sizeof(double);
    if (smem > c->lds_max) throw Fail(MYQL_E_UNSUPPORTED, "staging area exceeds the device's LDS per block");
    const unsigned nb = static_cast<unsigned>((n_paths + DBG_PATH_BLK - 1) / DBG_PATH_BLK);
    double *o = dout.as<double>();
    hipLaunchKernelGGL(debug_path_payoffs_kernel, dim3(nb), dim3(DBG_PATH_BLK), smem, c->stream, dj.as<Job<double>>(),
                       dx.as<double>(), static_cast<int>(n_paths), static_cast<int>(n_steps), o, o + no, o + 2 * no);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(base, o, no * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(up, o + no, no * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(dn, o + 2 * no, no * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_normcdfinv(myql_ctx *c, const double *p, int64_t n, double *out64, float *out32) {
  if (!p || !out64 || !out32) return MYQL_E_INVALID_ARG;
  return debug_map(c, 0, p, nullptr, n, out64, out32);
}

int myql_debug_poissinv(myql_ctx *c, const double *lambda, const double *u, int64_t n, double *out) {
  if (!lambda || !u || !out) return MYQL_E_INVALID_ARG;
  return debug_map(c, 1, lambda, u, n, out, nullptr);
}

int myql_debug_poisson_counts(myql_ctx *c, const double *lambda, int32_t n_slots, const double *u, int64_t n,
                              int32_t *out) {
  if (!c || !lambda || !u || !out || (n_slots != 1 && n_slots != 5)) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (n <= 0) return MYQL_OK;
    DevBuf dl, du, dout;
    dl.reserve(n * n_slots * sizeof(double));
    du.reserve(n * sizeof(double));
    dout.reserve(n * n_slots * sizeof(int32_t));
    HIPCHECK(hipMemcpyAsync(dl.p, lambda, n * n_slots * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(du.p, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const unsigned nb = static_cast<unsigned>((n + BLOCK - 1) / BLOCK);
    if (n_slots == 1)
      hipLaunchKernelGGL(debug_poisson_counts_kernel<1>, dim3(nb), dim3(BLOCK), 0, c->stream, dl.as<double>(),
                         du.as<double>(), n, dout.as<int32_t>());
    else
      hipLaunchKernelGGL(debug_poisson_counts_kernel<5>, dim3(nb), dim3(BLOCK), 0, c->stream, dl.as<double>(),
                         du.as<double>(), n, dout.as<int32_t>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, dout.p, n * n_slots * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_qmc_points(myql_ctx *c, uint64_t seed, uint64_t replicate, int32_t dims, uint64_t n_points,
                          uint64_t n_streams, uint32_t *out) {
  if (!c || !out || dims < 1 || dims > QMC_DMAX || n_points == 0 || n_streams == 0 || n_streams % 64 != 0 ||
      n_points > (1ull << 24))
    return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    qmc_tables(c);
    std::vector<uint32_t> sh(static_cast<size_t>(dims));
    qmc_shift(seed, replicate, dims, sh.data());
    DevBuf dsh, dout;
    dsh.reserve(sh.size() * sizeof(uint32_t));
    dout.reserve(n_points * static_cast<size_t>(dims) * sizeof(uint32_t));
    HIPCHECK(hipMemcpyAsync(dsh.p, sh.data(), sh.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    const uint64_t n_waves = n_streams / 64, n_it = (n_points + n_streams - 1) / n_streams;
    const unsigned nb = static_cast<unsigned>((n_waves + BLOCK / 64 - 1) / (BLOCK / 64));
    hipLaunchKernelGGL(debug_qmc_kernel, dim3(nb), dim3(BLOCK), 0, c->stream, c->qmc_V.as<uint32_t>(),
                       c->qmc_lo.as<uint32_t>(), dsh.as<uint32_t>(), static_cast<int>(dims), n_waves, n_it, n_points,
                       dout.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, dout.p, n_points * static_cast<size_t>(dims) * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_set_launch_shape(myql_ctx *c, int32_t precision, int32_t n_factors, int32_t lockstep_width,
                                int32_t path_dependent, int32_t block, int32_t unroll, int32_t grid_factor,
                                int32_t four_wave) {
  if (!c) return MYQL_E_INVALID_ARG;
  const bool pd = path_dependent != 0;
  if ((precision != MYQL_FP32 && precision != MYQL_FP64) || grid_factor < 1 ||
      !myql::mc_supported(n_factors, lockstep_width, pd)) {
    c->err = "bad launch shape arguments";
    return MYQL_E_INVALID_ARG;
  }
  bool ok = false;
  for (int si = 0; si < myql::mc_tune_shape_count(n_factors, lockstep_width, pd); ++si)
    ok = ok || (myql::MC_TUNE_SHAPES[si].blk == block && myql::MC_TUNE_SHAPES[si].unr == unroll &&
                myql::MC_TUNE_SHAPES[si].occ == (four_wave ? 4 : 0));
  if (!ok) {
    c->err = "launch shape not compiled for this kernel";
    return MYQL_E_UNSUPPORTED;
  }
  c->tuned[std::make_tuple(static_cast<int>(precision), static_cast<int>(n_factors),
                           static_cast<int>(lockstep_width), pd)] = {{block, unroll, four_wave ? 4 : 0}, grid_factor,
                                                                     "set"};
  c->err.clear();
  return MYQL_OK;
}

int myql_debug_chi2_quantile(myql_ctx *c, const double *dof, const double *p, int64_t n, double *out) {
  if (!dof || !p || !out) return MYQL_E_INVALID_ARG;
  return debug_map(c, 2, dof, p, n, out, nullptr);
}

int myql_debug_cv_jump_expectation(myql_ctx *c, const myql_model *model, const myql_instrument *instrument,
                                   int32_t greek_mode, double S0, const double *spots, int64_t n, double *tab,
                                   double *dir, double *ddir, int32_t *table_on) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!model || !instrument || !spots || !tab || !dir || !ddir || !table_on || n < 0 || !(S0 > 0.0))
      throw Fail(MYQL_E_INVALID_ARG, "null argument, negative count or non-positive S0");
    if (model->jump_kind != MYQL_JUMP_MERTON && model->jump_kind != MYQL_JUMP_KOU)
      throw Fail(MYQL_E_INVALID_ARG, "model has no jumps");
    if (instrument->n_strikes < 1 || instrument->n_strikes > 64 || !instrument->strikes || !(instrument->maturity > 0.0))
      throw Fail(MYQL_E_INVALID_ARG, "instrument needs 1..64 strikes and a positive maturity");
    for (int64_t i = 0; i < n; ++i)
      if (!(spots[i] > 0.0)) throw Fail(MYQL_E_INVALID_ARG, "spots must be positive");
    const int nk = static_cast<int>(instrument->n_strikes);
    CvSet S{};
    DevBuf ks, d_sp, d_out;
    ks.reserve(nk * sizeof(double));
    HIPCHECK(hipMemcpyAsync(ks.p, instrument->strikes, nk * sizeof(double), hipMemcpyHostToDevice, c->stream));
    S.strikes = ks.as<double>();
    S.nk = nk;
    S.payoff = instrument->payoff;
    S.call = instrument->option_type == MYQL_CALL ? 1 : 0;
    S.smooth = (greek_mode != MYQL_GREEK_NONE && instrument->payoff != MYQL_PAYOFF_VANILLA) ? 1 : 0;
    if (S.smooth && !(instrument->payoff_eps > 0.0))
      throw Fail(MYQL_E_INVALID_ARG, "a digital payoff in a Greek mode needs payoff_eps > 0 here");
    S.eps_scale = S.smooth ? 0.70710678118654752440 / instrument->payoff_eps : 0.0;
    cv_terminal_jump(c, *model, instrument->maturity, S0, instrument->strikes, S);
    *table_on = S.rb_on;
    if (S.rb_on && n > 0) {
      const size_t m = static_cast<size_t>(n) * nk;
      d_sp.reserve(n * sizeof(double));
      d_out.reserve(3 * m * sizeof(double));
      HIPCHECK(hipMemcpyAsync(d_sp.p, spots, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      double *o = d_out.as<double>();
      HIPCHECK(launch_cv_debug_rb(S, d_sp.as<double>(), n, o, o + m, o + 2 * m, c->stream));
      HIPCHECK(hipMemcpyAsync(tab, o, m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipMemcpyAsync(dir, o + m, m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipMemcpyAsync(ddir, o + 2 * m, m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_cv_train(myql_ctx *c, const myql_mc_job *jobs, int64_t n_jobs, const myql_cv_config *cfg, myql_cv_net *net) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!net) throw Fail(MYQL_E_INVALID_ARG, "null net");
    cv_validate(jobs, n_jobs, cfg);
    cv_train(c, jobs, n_jobs, cfg, net);
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_mc_calculate_batch_cv(myql_ctx *c, const myql_mc_job *jobs, int64_t n_jobs, const myql_cv_config *cfg,
                               const myql_cv_net *net, myql_cv_result *cv_results) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!net) throw Fail(MYQL_E_INVALID_ARG, "null net");
    cv_validate(jobs, n_jobs, cfg);
    cv_price(c, jobs, n_jobs, cfg, net, cv_results);
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_cv_input_grad(myql_ctx *c, const float *params, const float *feats, int64_t n, float *out) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!params || !feats || !out || n < 1) throw Fail(MYQL_E_INVALID_ARG, "bad arguments");
    DevBuf p, f, o;
    p.reserve(CV_NPARAM * sizeof(float));
    f.reserve(n * CV_IN * sizeof(float));
    o.reserve(n * CV_OUT * 3 * sizeof(float));
    HIPCHECK(hipMemcpyAsync(p.p, params, CV_NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(f.p, feats, n * CV_IN * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(launch_cv_grad_debug(p.as<float>(), f.as<float>(), static_cast<int>(n), o.as<float>(), c->stream));
    HIPCHECK(hipMemcpyAsync(out, o.p, n * CV_OUT * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

int myql_debug_cv_param_grad(myql_ctx *c, const float *params, const float *feats, const float *targets, int64_t n,
                             float *grad, double *loss) {
  if (!c) return MYQL_E_INVALID_ARG;
  try {
    set_device(c);
    if (!params || !feats || !targets || !grad || n < CV_TROWS || n % CV_TROWS != 0)
      throw Fail(MYQL_E_INVALID_ARG, "bad arguments (n must be a positive multiple of 32)");
    const int nb = static_cast<int>(n / CV_TROWS);
    DevBuf p, f, t, gp, lp, stp;
    p.reserve(CV_NPARAM * sizeof(float));
    f.reserve(n * CV_IN * sizeof(float));
    t.reserve(n * CV_OUT * sizeof(float));
    gp.reserve(static_cast<size_t>(nb) * CV_NPARAM * sizeof(float));
    lp.reserve(nb * sizeof(float));
    stp.reserve(sizeof(int));
    HIPCHECK(hipMemcpyAsync(p.p, params, CV_NPARAM * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(f.p, feats, n * CV_IN * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(t.p, targets, n * CV_OUT * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemsetAsync(stp.p, 0, sizeof(int), c->stream));
    CvTrainArgs a{p.as<float>(), nullptr, nullptr, gp.as<float>(), lp.as<float>(), nullptr, nullptr, 0, CvNorm{},
                  0, stp.as<int>(), static_cast<int>(n), CV_OUT, 0.f, 1.f, false, f.as<float>(), t.as<float>()};
    HIPCHECK(launch_cv_train_step(a, c->stream));
    std::vector<float> g(static_cast<size_t>(nb) * CV_NPARAM), l(nb);
    HIPCHECK(hipMemcpyAsync(g.data(), gp.p, g.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(l.data(), lp.p, nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    double ls = 0.0;
    for (int b = 0; b < nb; ++b) ls += l[b];
    for (int i = 0; i < CV_NPARAM; ++i) {
      double s = 0.0;
      for (int b = 0; b < nb; ++b) s += g[static_cast<size_t>(b) * CV_NPARAM + i];
      grad[i] = static_cast<float>(s);
    }
    if (loss) *loss = ls / (static_cast<double>(n) * CV_OUT);
    c->err.clear();
    return MYQL_OK;
  } catch (const Fail &e) {
    return fail(c, e);
  }
}

}  // extern "C"
