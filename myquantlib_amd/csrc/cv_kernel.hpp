// Learned martingale control variate (no counterpart in the reference, whose
// README.md:332 lists control variates as future work).
//
// u(t, x) ~ E[e^{-r(T-t)} P(S_T) | X_t = x] is an MLP 6 -> 64 -> 64 -> 64 -> 1
// with SiLU activations (8,833 parameters) on the features
//   f = (t, log(S/K), v1, v2, T, log(K/S0)) (each scaled, see CvNorm),
// trained by regression of the discounted payoff on the state at a few
// control points of engine-generated paths. The control variate of one path
// and strike is the discrete stochastic integral
//   CV = sum_j e^{r(T - t_j)} grad_x u(t_j, X_j) . sigma(X_j) dW_j
// with the integrand frozen at the control point t_j and dW_j the sum of the
// stepper's own Brownian increments over [t_j, t_{j+1}) (sqrt(dt) G for the
// stock's independent part, sqrt(dt) Z for each NV variance factor). Every
// term has zero mean whatever the weights, so P - CV is an unbiased payoff
// with lower variance when u is close to the price function.
//
// Kernels:
//   cv_paths_kernel   the engine's base path (same streams, partition and
//                     draws as mc_paths_kernel) recording state and
//                     increments at the control points and the terminal spot
//   cv_eval_kernel    per (path, 32-strike tile, control point): MLP forward
//                     and input gradient on BF16 MFMA (FP32 accumulate), then
//                     per-strike FP64 sums of P - CV and P
//   cv_train_kernel   FP32 forward/backward of 32 sampled rows per block,
//                     per-block parameter gradients
//   cv_adam_kernel    gradient reduction (fixed order) and Adam update
#pragma once
#include "cv_types.hpp"
#include "mlmc_kernel.hpp"

namespace myql {

__device__ __forceinline__ short bf16_bits(float x) {
  const __bf16 b = static_cast<__bf16>(x);
  return __builtin_bit_cast(short, b);
}

typedef short cv_v4s __attribute__((ext_vector_type(4)));
typedef float cv_v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float sigmoidf(float a) { return 1.0f / (1.0f + __expf(-a)); }

// ---------------------------------------------------------------------------
// recording path kernel
// ---------------------------------------------------------------------------
template <typename R, int NF>
__global__ void __launch_bounds__(BLOCK) cv_paths_kernel(const Job<R> *__restrict__ J, CvSet set) {
  constexpr int NFX = NF > 0 ? NF : 1;
  const uint64_t S = J->S, M = J->M;
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * BLOCK + threadIdx.x;
  if (s >= S) return;
  Rng g;
  g.s0 = J->streams[4 * s + 0];
  g.s1 = J->streams[4 * s + 1];
  g.s2 = J->streams[4 * s + 2];
  g.s3 = J->streams[4 * s + 3];
  const uint64_t q = M / S, rem = M % S;
  const uint64_t count = q + (s < rem ? 1 : 0);
  const uint64_t first = s * q + (s < rem ? s : rem);
  const VarParams<R> *__restrict__ P = &J->var[V_BASE];
  const VarParams<R> *__restrict__ PP[1] = {P};
  const R sqdt = static_cast<R>(set.sqdt);
  const unsigned steps = static_cast<unsigned>(J->steps);
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t p = first + i;
    R x = static_cast<R>(J->logS0);
    R v[NFX];
#pragma unroll
    for (int f = 0; f < NFX; ++f) v[f] = P->v0[f];
    int j = 0;
    R dwp = 0, dwv[2] = {0, 0};
    if constexpr (NF == 0) {
      // zero-factor European: one step to maturity (ASVJstepper.hpp:145-161),
      // one control point at t = 0 with dW = sqrt(T) G (set.sqdt = sqrt(T))
      set.rec[0 * M + p] = static_cast<float>(x);
      set.rec[1 * M + p] = 0.0f;
      set.rec[2 * M + p] = 0.0f;
      const R G = normcdfinv<R>(uniform(g));
      x += P->gdrift_n + P->gvol_n * G;
      dwp = sqdt * G;
    } else {
    for (unsigned step = 0; step < steps; ++step) {
      if (j < set.nj && step == static_cast<unsigned>(set.step_of[j])) {
        if (j > 0) {  // close the previous block
          set.rec[((j - 1) * CV_REC + 3) * M + p] = static_cast<float>(dwp);
          set.rec[((j - 1) * CV_REC + 4) * M + p] = static_cast<float>(dwv[0]);
          set.rec[((j - 1) * CV_REC + 5) * M + p] = static_cast<float>(dwv[1]);
        }
        set.rec[(j * CV_REC + 0) * M + p] = static_cast<float>(x);
        set.rec[(j * CV_REC + 1) * M + p] = NF > 0 ? static_cast<float>(v[0]) : 0.0f;
        set.rec[(j * CV_REC + 2) * M + p] = NF > 1 ? static_cast<float>(v[NFX - 1]) : 0.0f;
        dwp = 0;
        dwv[0] = dwv[1] = 0;
        ++j;
      }
      // one step of ASVJStepper::multi_step with the draws the engine makes
      const R G = normcdfinv<R>(uniform(g));
      dwp += sqdt * G;
      if constexpr (NF == 0) {
        x += P->gdrift + P->gvol * G;
      } else {
        R vn[NFX];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          if (J->scheme[f] == 1) {
            const R Z = normcdfinv<R>(uniform(g));
            vn[f] = nv_step<R, NF>(P, f, v[f], Z);
            dwv[f] += sqdt * Z;
          } else {
            vn[f] = evolve_one<R, NFX>(J, P, f, v[f], g);
          }
        }
        R a = 0, d = 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          a += P->C[f][3] * (v[f] + vn[f]);
          d += P->C[f][0] + P->C[f][1] * v[f] + P->C[f][2] * vn[f];
        }
        x += d + dsqrt<R>(a) * G;
#pragma unroll
        for (int f = 0; f < NF; ++f) v[f] = vn[f];
      }
    }
    }
    set.rec[((set.nj - 1) * CV_REC + 3) * M + p] = static_cast<float>(dwp);
    set.rec[((set.nj - 1) * CV_REC + 4) * M + p] = static_cast<float>(dwv[0]);
    set.rec[((set.nj - 1) * CV_REC + 5) * M + p] = static_cast<float>(dwv[1]);
    R xs[1] = {x};
    euro_jump<R, 1>(J, PP, xs, g);
    set.ST[p] = static_cast<double>(dexp<R>(xs[0]));
    // the compensated jump factor of the European's single terminal jump
    if (set.dJ) set.dJ[p] = static_cast<float>(expm1(static_cast<double>(xs[0]) - static_cast<double>(x)));
  }
}

// ---------------------------------------------------------------------------
// MLP weights in LDS for the MFMA evaluation
// ---------------------------------------------------------------------------
struct CvLds {
  short w2[CV_H * CV_PITCH], w3[CV_H * CV_PITCH];    // [n][k]
  short w2t[CV_H * CV_PITCH], w3t[CV_H * CV_PITCH];  // [k][n]
  float w1[CV_H * CV_IN], b1[CV_H], b2[CV_H], b3[CV_H], w4[CV_H];
};

__device__ __forceinline__ void cv_load_lds(CvLds &L, const float *__restrict__ prm) {
  for (int i = threadIdx.x; i < CV_H * CV_H; i += blockDim.x) {
    const int n = i / CV_H, k = i % CV_H;
    const short a = bf16_bits(prm[CV_W2 + i]), b = bf16_bits(prm[CV_W3 + i]);
    L.w2[n * CV_PITCH + k] = a;
    L.w3[n * CV_PITCH + k] = b;
    L.w2t[k * CV_PITCH + n] = a;
    L.w3t[k * CV_PITCH + n] = b;
  }
  for (int i = threadIdx.x; i < CV_H * CV_IN; i += blockDim.x) L.w1[i] = prm[CV_W1 + i];
  for (int i = threadIdx.x; i < CV_H; i += blockDim.x) {
    L.b1[i] = prm[CV_B1 + i];
    L.b2[i] = prm[CV_B2 + i];
    L.b3[i] = prm[CV_B3 + i];
    L.w4[i] = prm[CV_W4 + i];
  }
}

// Feature index n held by this lane in accumulator element a of tile nt
// (v_mfma_f32_32x32x8bf16_1k: D[8(a/4) + 4(lane/32) + a%4][lane%32]).
__device__ __forceinline__ int cv_feat(int nt, int a, int hi) { return 32 * nt + 8 * (a >> 2) + 4 * hi + (a & 3); }

// A zero the compiler cannot see through: LDS weight reads indexed with it
// are re-issued on every call instead of being hoisted out of the caller's
// path loop into registers (hundreds of VGPRs for the MFMA A fragments).
__device__ __forceinline__ int cv_opaque_zero() {
  int z = 0;
  asm volatile("" : "+s"(z));
  return z;
}

// out^T = W x^T for 64 outputs: A = W rows (LDS), B = x^T fragments.
__device__ __forceinline__ void cv_layer(const short *__restrict__ W, const cv_v4s (&B)[8], cv_v16f (&acc)[2],
                                         int lane, int z) {
  const int r = lane & 31, hi = lane >> 5;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    cv_v16f c = {0};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const cv_v4s a = *reinterpret_cast<const cv_v4s *>(W + (32 * nt + r) * CV_PITCH + 8 * s + 4 * hi + z);
      c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, B[s], c, 0, 0, 0);
    }
    acc[nt] = c;
  }
}

// Input gradient d net / d f_i for i = 1, 2, 3 (log-moneyness, v1, v2) of the
// row held by lanes (lane % 32, lane % 32 + 32), returned in both lanes.
// Activations go straight into the BF16 B fragments of the next product
// (fragment s = features 8s + 4(lane/32) + 0..3, the accumulator layout of
// tile s/4); SiLU derivatives are kept as FP16.
__device__ __forceinline__ void cv_grad(const CvLds &L, const float (&f)[CV_IN], float &g1, float &g2, float &g3,
                                        int lane) {
  const int hi = lane >> 5;
  const int z = cv_opaque_zero();
  _Float16 d1[2][16], d2[2][16];
  cv_v4s B[8];
  // layer 1 on VALU (K = 6)
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int n = cv_feat(nt, a, hi) + z;
      float y = L.b1[n];
#pragma unroll
      for (int i = 0; i < CV_IN; ++i) y += L.w1[n * CV_IN + i] * f[i];
      const float sg = sigmoidf(y), h = y * sg;
      d1[nt][a] = static_cast<_Float16>(sg + h * (1.0f - sg));
      B[4 * nt + (a >> 2)][a & 3] = bf16_bits(h);
    }
  cv_v16f acc[2];
  cv_layer(L.w2, B, acc, lane, z);
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float y = acc[nt][a] + L.b2[cv_feat(nt, a, hi) + z];
      const float sg = sigmoidf(y), h = y * sg;
      d2[nt][a] = static_cast<_Float16>(sg + h * (1.0f - sg));
      B[4 * nt + (a >> 2)][a & 3] = bf16_bits(h);
    }
  cv_layer(L.w3, B, acc, lane, z);
  // delta3 = w4 * silu'(a3)
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int n = cv_feat(nt, a, hi) + z;
      const float y = acc[nt][a] + L.b3[n];
      const float sg = sigmoidf(y), h = y * sg;
      B[4 * nt + (a >> 2)][a & 3] = bf16_bits(L.w4[n] * (sg + h * (1.0f - sg)));
    }
  cv_layer(L.w3t, B, acc, lane, z);  // W3^T delta3
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int a = 0; a < 16; ++a) B[4 * nt + (a >> 2)][a & 3] = bf16_bits(acc[nt][a] * static_cast<float>(d2[nt][a]));
  cv_layer(L.w2t, B, acc, lane, z);  // W2^T delta2
  float s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int n = cv_feat(nt, a, hi) + z;
      const float d = acc[nt][a] * static_cast<float>(d1[nt][a]);
      s1 += L.w1[n * CV_IN + 1] * d;
      s2 += L.w1[n * CV_IN + 2] * d;
      s3 += L.w1[n * CV_IN + 3] * d;
    }
  g1 = s1 + __shfl_xor(s1, 32, 64);
  g2 = s2 + __shfl_xor(s2, 32, 64);
  g3 = s3 + __shfl_xor(s3, 32, 64);
}

__device__ __forceinline__ double cv_payoff(int payoff, int call, double S, double K) {
  const double diff = call ? S - K : K - S;
  if (payoff == 0) return diff > 0.0 ? diff : 0.0;
  const double w = diff > 0.0 ? 1.0 : 0.0;
  return payoff == 1 ? w : S * w;
}

// ---------------------------------------------------------------------------
// evaluation: each wave takes paths p = wave, wave + n_waves, ...; lanes
// (r, r + 32) hold row r = strike 32*tile + r. Per-wave partial sums
// part[wave][k][4] = (sum P-CV, sum (P-CV)^2, sum P, sum P^2).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2))) cv_eval_kernel(const float *__restrict__ prm, CvNorm nm, CvSet set,
                                                        double *__restrict__ part) {
  __shared__ CvLds L;
  cv_load_lds(L, prm);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r = lane & 31;
  const uint64_t n_waves = static_cast<uint64_t>(gridDim.x) * (BLOCK / 64);
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (BLOCK / 64) + threadIdx.x / 64;
  const uint64_t M = set.M;
  const int n_tiles = (set.nk + 31) / 32;
  // strike tiles outermost (wave-uniform trip count); the path records are
  // read once per tile
  for (int t = 0; t < n_tiles; ++t) {
    const int k = 32 * t + r;
    const bool live = k < set.nk;
    const double K = live ? set.strikes[k] : static_cast<double>(nm.S0);
    const float lk = static_cast<float>(log(K));
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (uint64_t p = wave; p < M; p += n_waves) {
      double cv = 0.0;
      float ux_last = 0.f;
      for (int j = 0; j < set.nj; ++j) {
        const float x = set.rec[(j * CV_REC + 0) * M + p];
        const float v1 = set.rec[(j * CV_REC + 1) * M + p];
        const float v2 = set.rec[(j * CV_REC + 2) * M + p];
        const float dwp = set.rec[(j * CV_REC + 3) * M + p];
        const float dw1 = set.rec[(j * CV_REC + 4) * M + p];
        const float dw2 = set.rec[(j * CV_REC + 5) * M + p];
        const float sv1 = sqrtf(fmaxf(v1, 0.f)), sv2 = sqrtf(fmaxf(v2, 0.f));
        // stock and variance noise of the frozen integrand
        float sx_perp, sx1, sx2;
        if (set.nf == 0) {
          sx_perp = set.vol;
          sx1 = sx2 = 0.f;
        } else {
          const float vp = set.rperp[0] * fmaxf(v1, 0.f) + (set.nf > 1 ? set.rperp[1] * fmaxf(v2, 0.f) : 0.f);
          sx_perp = sqrtf(vp);
          sx1 = set.rho[0] * sv1;
          sx2 = set.nf > 1 ? set.rho[1] * sv2 : 0.f;
        }
        const float ex = sx_perp * dwp + sx1 * dw1 + sx2 * dw2;  // stock martingale increment
        const float e1 = set.sig[0] * sv1 * dw1, e2 = set.sig[1] * sv2 * dw2;
        float f[CV_IN];
        f[0] = set.t_of[j] * nm.ts;
        f[1] = (x - lk) * nm.xs;
        f[2] = v1 * nm.vs0;
        f[3] = v2 * nm.vs1;
        f[4] = static_cast<float>(set.T) * nm.ts;
        f[5] = (lk - nm.logS0) * nm.xs;
        float g1, g2, g3;
        cv_grad(L, f, g1, g2, g3, lane);
        const float ux = nm.S0 * nm.xs * g1, u1 = nm.S0 * nm.vs0 * g2, u2 = nm.S0 * nm.vs1 * g3;
        cv += set.disc_of[j] * static_cast<double>(ux * ex + u1 * e1 + u2 * e2);
        ux_last = ux;
      }
      // terminal jump: u_x dS/S at the last control point times the
      // compensated jump factor (independent of the path, mean 0)
      if (set.dJ) cv += set.disc_of[set.nj - 1] * static_cast<double>(ux_last * set.dJ[p]);
      if (live) {
        const double P = cv_payoff(set.payoff, set.call, set.ST[p], K);
        const double Pa = P - cv;
        a0 += Pa;
        a1 += Pa * Pa;
        a2 += P;
        a3 += P * P;
      }
    }
    if (lane < 32 && live) {
      double *o = part + (wave * 64 + k) * 4;
      o[0] = a0;
      o[1] = a1;
      o[2] = a2;
      o[3] = a3;
    }
  }
}

// sums[k][c] = sum over waves of part[w][k][c], in wave order
__global__ void cv_reduce_kernel(const double *__restrict__ part, int n_waves, int nk, double *__restrict__ sums) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk * 4) return;
  const int k = i / 4, c = i % 4;
  double s = 0.0;
  for (int w = 0; w < n_waves; ++w) s += part[(static_cast<size_t>(w) * 64 + k) * 4 + c];
  sums[i] = s;
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
  float w2[CV_H][CV_H + 1], w3[CV_H][CV_H + 1];
  float w1[CV_H * CV_IN], b1[CV_H], b2[CV_H], b3[CV_H], w4[CV_H];
  float f[CV_TROWS][CV_IN];
  float h1[CV_TROWS][CV_H + 1], h2[CV_TROWS][CV_H + 1], h3[CV_TROWS][CV_H + 1];
  float d1[CV_TROWS][CV_H + 1], d2[CV_TROWS][CV_H + 1], d3[CV_TROWS][CV_H + 1];
  float ypart[8][CV_TROWS];
  float dy[CV_TROWS], target[CV_TROWS], loss[CV_TROWS];
};

// Rows: a random (set, path, control point, strike) per row of each
// minibatch, from a counter hash of (seed, step, row). Feature and target
// construction matches cv_eval_kernel. Writes the block's parameter
// gradient of mean (u - target)^2 over the whole minibatch, and its loss.
__device__ __forceinline__ void cv_row(const CvSet *__restrict__ sets, int n_sets, CvNorm nm, uint64_t h,
                                       float (&f)[CV_IN], float &target) {
  const CvSet &S = sets[h % static_cast<uint64_t>(n_sets)];
  const uint64_t h2 = cv_hash(h);
  const uint64_t p = h2 % S.M;
  const int j = static_cast<int>((h2 >> 40) % static_cast<uint64_t>(S.nj));
  const int k = static_cast<int>(cv_hash(h2) % static_cast<uint64_t>(S.nk));
  const double K = S.strikes[k];
  const float lk = static_cast<float>(log(K));
  f[0] = S.t_of[j] * nm.ts;
  f[1] = (S.rec[(j * CV_REC + 0) * S.M + p] - lk) * nm.xs;
  f[2] = S.rec[(j * CV_REC + 1) * S.M + p] * nm.vs0;
  f[3] = S.rec[(j * CV_REC + 2) * S.M + p] * nm.vs1;
  f[4] = static_cast<float>(S.T) * nm.ts;
  f[5] = (lk - nm.logS0) * nm.xs;
  target = static_cast<float>(cv_payoff(S.payoff, S.call, S.ST[p], K) / S.disc_of[j] * S.inv_S0);
}

__global__ void __launch_bounds__(BLOCK) cv_train_kernel(const float *__restrict__ prm, const CvSet *__restrict__ sets,
                                                         int n_sets, CvNorm nm, uint64_t seed,
                                                         const int *__restrict__ step, int batch,
                                                         float *__restrict__ gpart, float *__restrict__ lpart,
                                                         const float *__restrict__ fixed_f,
                                                         const float *__restrict__ fixed_t) {
  __shared__ CvTrainLds L;
  const int tid = threadIdx.x;
  const int r = tid % CV_TROWS, q8 = tid / CV_TROWS;  // row, slice of 8 outputs
  for (int i = tid; i < CV_H * CV_H; i += BLOCK) {
    L.w2[i / CV_H][i % CV_H] = prm[CV_W2 + i];
    L.w3[i / CV_H][i % CV_H] = prm[CV_W3 + i];
  }
  for (int i = tid; i < CV_H * CV_IN; i += BLOCK) L.w1[i] = prm[CV_W1 + i];
  for (int i = tid; i < CV_H; i += BLOCK) {
    L.b1[i] = prm[CV_B1 + i];
    L.b2[i] = prm[CV_B2 + i];
    L.b3[i] = prm[CV_B3 + i];
    L.w4[i] = prm[CV_W4 + i];
  }
  const int row = blockIdx.x * CV_TROWS + r;
  if (q8 == 0) {
    float f[CV_IN], tg;
    if (fixed_f) {  // caller-given rows (gradient checks)
      for (int i = 0; i < CV_IN; ++i) f[i] = fixed_f[row * CV_IN + i];
      tg = fixed_t[row];
    } else {
      cv_row(sets, n_sets, nm, cv_hash(seed ^ cv_hash(static_cast<uint64_t>(*step) * 0x100000001ull + row)), f, tg);
    }
    for (int i = 0; i < CV_IN; ++i) L.f[r][i] = f[i];
    L.target[r] = tg;
  }
  __syncthreads();
  const float inv_b = 1.0f / static_cast<float>(batch);
  // forward
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = L.b1[n];
    for (int i = 0; i < CV_IN; ++i) z += L.w1[n * CV_IN + i] * L.f[r][i];
    const float sg = sigmoidf(z), h = z * sg;
    L.h1[r][n] = h;
    L.d1[r][n] = sg + h * (1.0f - sg);
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = L.b2[n];
    for (int k = 0; k < CV_H; ++k) z += L.w2[n][k] * L.h1[r][k];
    const float sg = sigmoidf(z), h = z * sg;
    L.h2[r][n] = h;
    L.d2[r][n] = sg + h * (1.0f - sg);
  }
  __syncthreads();
  float yp = 0.f;
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    float z = L.b3[n];
    for (int k = 0; k < CV_H; ++k) z += L.w3[n][k] * L.h2[r][k];
    const float sg = sigmoidf(z), h = z * sg;
    L.h3[r][n] = h;
    L.d3[r][n] = sg + h * (1.0f - sg);
    yp += L.w4[n] * h;
  }
  L.ypart[q8][r] = yp;
  __syncthreads();
  if (q8 == 0) {
    float y = prm[CV_B4];
    for (int i = 0; i < 8; ++i) y += L.ypart[i][r];
    const float e = y - L.target[r];
    L.loss[r] = e * e;
    L.dy[r] = 2.0f * e * inv_b;
  }
  __syncthreads();
  // backward: deltas overwrite the activation derivatives
  for (int a = 0; a < 8; ++a) {
    const int n = q8 * 8 + a;
    L.d3[r][n] *= L.dy[r] * L.w4[n];
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int k = q8 * 8 + a;
    float s = 0.f;
    for (int n = 0; n < CV_H; ++n) s += L.w3[n][k] * L.d3[r][n];
    L.d2[r][k] *= s;
  }
  __syncthreads();
  for (int a = 0; a < 8; ++a) {
    const int k = q8 * 8 + a;
    float s = 0.f;
    for (int n = 0; n < CV_H; ++n) s += L.w2[n][k] * L.d2[r][n];
    L.d1[r][k] *= s;
  }
  __syncthreads();
  // parameter gradients of this block's rows
  float *g = gpart + static_cast<size_t>(blockIdx.x) * CV_NPARAM;
  for (int e = tid; e < CV_H * CV_H; e += BLOCK) {
    const int n = e / CV_H, k = e % CV_H;
    float s2 = 0.f, s3 = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr) {
      s2 += L.d2[rr][n] * L.h1[rr][k];
      s3 += L.d3[rr][n] * L.h2[rr][k];
    }
    g[CV_W2 + e] = s2;
    g[CV_W3 + e] = s3;
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
      s4 += L.dy[rr] * L.h3[rr][n];
    }
    g[CV_B1 + n] = s1;
    g[CV_B2 + n] = s2;
    g[CV_B3 + n] = s3;
    g[CV_W4 + n] = s4;
  }
  if (tid == 0) {
    float s = 0.f, l = 0.f;
    for (int rr = 0; rr < CV_TROWS; ++rr) {
      s += L.dy[rr];
      l += L.loss[rr];
    }
    g[CV_B4] = s;
    lpart[blockIdx.x] = l;
  }
}

// Adam on the block-summed gradient (fixed block order). state: m[NPARAM],
// v[NPARAM]; step counter read here, advanced by cv_step_kernel.
__global__ void cv_adam_kernel(float *__restrict__ prm, float *__restrict__ m, float *__restrict__ v,
                               const float *__restrict__ gpart, int n_blocks, const int *__restrict__ step, float lr,
                               const float *__restrict__ lpart, float *__restrict__ loss_hist, int batch) {
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
    loss_hist[t - 1] = l / static_cast<float>(batch);
  }
}

__global__ void cv_step_kernel(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// debug: BF16 MFMA input gradient of n rows of features (n multiple of 32
// handled per wave; rows beyond n are padded with zeros)
__global__ void __launch_bounds__(BLOCK) cv_grad_debug_kernel(const float *__restrict__ prm,
                                                              const float *__restrict__ feats, int n,
                                                              float *__restrict__ out) {
  __shared__ CvLds L;
  cv_load_lds(L, prm);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64;
  const int row = wave * 32 + (lane & 31);
  float f[CV_IN];
  for (int i = 0; i < CV_IN; ++i) f[i] = row < n ? feats[row * CV_IN + i] : 0.f;
  float g1, g2, g3;
  cv_grad(L, f, g1, g2, g3, lane);
  if (lane < 32 && row < n) {
    out[row * 3 + 0] = g1;
    out[row * 3 + 1] = g2;
    out[row * 3 + 2] = g3;
  }
}

}  // namespace myql
