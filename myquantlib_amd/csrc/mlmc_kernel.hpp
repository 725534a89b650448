// Multilevel Monte Carlo level kernels (reference
// MultiLevelMonteCarloPricer.hpp:244-341 evaluate_level, and
// CoupledASVJStepper.hpp:170-517 multi_step_coupled).
//
// One lane owns one stream and keeps it across levels and iterations of the
// Giles loop (the reference's per-thread RNG pool, :55-61): the kernel loads
// the stream state, prices its share of the level's new paths and writes the
// state back. Level 0 runs the plain stepper (Job::var[0]); level l > 0 runs
// fine (var[0], dt_f) and coarse (var[1], 2 dt_f) paths in one lane, the
// coarse path driven by the fine increments: G_c from the two fine stock
// normals weighted by their diffusion sizes, Z_c = (Z_1 + Z_2)/sqrt(2) for NV
// variance, and the fine variance itself for NCI/Exact.
#pragma once
#include "mc_kernel.hpp"

namespace myql {

template <typename R, int NF>
__device__ __forceinline__ R nv_step(const VarParams<R> *__restrict__ P, int f, R v, R Z) {
  const R root = dsqrt<R>(P->Kr[f][0] * v + P->Kr[f][1]) + P->Kr[f][2] * Z;
  return P->Kr[f][0] * (root * root) + P->Kr[f][1];
}

// single-slot variance step with the scheme's own draws
template <typename R, int NFX>
__device__ __forceinline__ R evolve_one(const Job<R> *__restrict__ J, const VarParams<R> *__restrict__ P, int f,
                                        R v, Rng &g, int *n_out = nullptr) {
  R va[1][NFX], vn[1][NFX];
  va[0][f] = v;
  const VarParams<R> *__restrict__ PP[1] = {P};
  evolve_factor<R, 1, NFX>(J, PP, f, va, vn, g, nullptr, nullptr, n_out);
  return vn[0][f];
}

template <typename R, int NF>
__device__ void sim_coupled(const Job<R> *__restrict__ J, Rng &g, R *__restrict__ stage, int tid) {
  constexpr int NFX = NF > 0 ? NF : 1;
  const VarParams<R> *__restrict__ F = &J->var[0];
  const VarParams<R> *__restrict__ C = &J->var[1];
  const R x0 = static_cast<R>(J->logS0);
  R xf = x0, xc = x0;
  R vf[NFX], vc[NFX];
#pragma unroll
  for (int f = 0; f < NFX; ++f) {
    vf[f] = F->v0[f];
    vc[f] = C->v0[f];
  }
  Trk<R> tf, tc;
  trk_init(J, tf, x0, F->dt);
  trk_init(J, tc, x0, C->dt);
  const bool pd = J->pathdep;
  const unsigned nc = static_cast<unsigned>(J->steps);
  const R SQ2 = R(1.4142135623730951);
  const VarParams<R> *__restrict__ FP[1] = {F};
  if (NF == 0 && !pd) {
    const R G = normcdfinv<R>(uniform(g));
    xf += F->gdrift_n + F->gvol_n * G;
    xc += C->gdrift_n + C->gvol_n * G;
    R jj[1] = {0};
    euro_jump<R, 1>(J, FP, jj, g);
    xf += jj[0];
    xc += jj[0];
  } else {
    if (!pd && NF > 0) {
      R jj[1] = {0};
      euro_jump<R, 1>(J, FP, jj, g);
      xf += jj[0];
      xc += jj[0];
    }
    for (unsigned k = 0; k < nc; ++k) {
      const R G1 = normcdfinv<R>(uniform(g));
      const R G2 = normcdfinv<R>(uniform(g));
      R Gc = (G1 + G2) / SQ2;
      R j1 = 0, j2 = 0;
      if constexpr (NF == 0) {
        xc += C->gdrift + C->gvol * Gc;
        xf += F->gdrift + F->gvol * G1;
        if (pd) {
          if (J->jump_kind) j1 = log_jump(J, F, g), xf += j1;
          trk_update(J, tf, xf, F->dt);
        }
        xf += F->gdrift + F->gvol * G2;
        if (pd) {
          if (J->jump_kind) j2 = log_jump(J, F, g), xf += j2;
          trk_update(J, tf, xf, F->dt);
          xc += j1 + j2;
          trk_update(J, tc, xc, C->dt);
        }
      } else {
        R mid[NF], nxt[NF], Z1[NF], Z2[NF], cn[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          Z1[f] = 0;
          if (J->scheme[f] == 1) {
            Z1[f] = normcdfinv<R>(uniform(g));
            mid[f] = nv_step<R, NF>(F, f, vf[f], Z1[f]);
          } else {
            mid[f] = evolve_one<R, NFX>(J, F, f, vf[f], g);
          }
        }
        R a1sq = 0, d1 = 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          a1sq += F->C[f][3] * (vf[f] + mid[f]);
          d1 += F->C[f][0] + F->C[f][1] * vf[f] + F->C[f][2] * mid[f];
        }
        const R a1 = dsqrt<R>(a1sq);
        xf += d1 + a1 * G1;
        if (pd) {
          if (J->jump_kind) j1 = log_jump(J, F, g), xf += j1;
          trk_update(J, tf, xf, F->dt);
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          Z2[f] = 0;
          if (J->scheme[f] == 1) {
            Z2[f] = normcdfinv<R>(uniform(g));
            nxt[f] = nv_step<R, NF>(F, f, mid[f], Z2[f]);
          } else {
            nxt[f] = evolve_one<R, NFX>(J, F, f, mid[f], g);
          }
        }
        R a2sq = 0, d2 = 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          a2sq += F->C[f][3] * (mid[f] + nxt[f]);
          d2 += F->C[f][0] + F->C[f][1] * mid[f] + F->C[f][2] * nxt[f];
        }
        const R a2 = dsqrt<R>(a2sq);
        xf += d2 + a2 * G2;
        if (pd) {
          if (J->jump_kind) j2 = log_jump(J, F, g), xf += j2;
          trk_update(J, tf, xf, F->dt);
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
          cn[f] = (J->scheme[f] == 1) ? nv_step<R, NF>(C, f, vc[f], (Z1[f] + Z2[f]) / SQ2) : nxt[f];
        if (a1 + a2 > R(1e-15)) Gc = (a1 * G1 + a2 * G2) / dsqrt<R>(a1 * a1 + a2 * a2);
        R acsq = 0, dc = 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          acsq += C->C[f][3] * (vc[f] + cn[f]);
          dc += C->C[f][0] + C->C[f][1] * vc[f] + C->C[f][2] * cn[f];
        }
        xc += dc + dsqrt<R>(acsq) * Gc;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          vf[f] = nxt[f];
          vc[f] = cn[f];
        }
        if (pd) {
          xc += j1 + j2;
          trk_update(J, tc, xc, C->dt);
        }
      }
    }
  }
  finalize_stage(J, tf, xf, F->dt, F->total_time, stage, 0, tid);
  finalize_stage(J, tc, xc, C->dt, C->total_time, stage, 1, tid);
}

// Price dN new paths of one level; per block the FP64 (sum, sum of squares)
// of the level's price difference and, with GK (Greek mode Essential or
// Full), of its delta and gamma differences: each path's spot-up and
// spot-down payoffs (the reference's calculate_to_buffer output,
// MultiLevelMonteCarloPricer.hpp:280,321-324) give the central differences
// (up - dn) / 2h and (up - 2 base + dn) / h^2 of its fine and its coarse
// path; block_sums holds MLMC_SUMS doubles per block.
// Compiled for four waves per SIMD: 128 VGPRs with 5 spilled for one factor
// (133 and three waves otherwise), 4 spilled for two (194 and two waves); at
// 262,144 streams the Heston up-and-out MLMC run takes 697 ms against 764 ms
// for the three-wave build at its best, 196,608 streams.
constexpr int MLMC_SUMS = 6;

template <typename R, int NF, bool GK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) mlmc_level_kernel(const Job<R> *__restrict__ J,
                                                           uint64_t *__restrict__ streams,
                                                           const uint64_t *__restrict__ dNp, int coupled,
                                                           double *__restrict__ block_sums) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double s_red[MLMC_SUMS * (BLOCK / 64)];
  R *stage = reinterpret_cast<R *>(smem);
  const int tid = threadIdx.x;
  const uint64_t S = J->S;
  const uint64_t dN = *dNp;  // new paths of this level (device memory: the launch is a replayed graph)
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * BLOCK + tid;
  double acc[MLMC_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (s < S) {
    Rng g = {streams[4 * s], streams[4 * s + 1], streams[4 * s + 2], streams[4 * s + 3]};
    const uint64_t q = dN / S, rem = dN % S;
    const uint64_t cnt = q + (s < rem ? 1 : 0);
    const double K = J->strikes[0];
    for (uint64_t i = 0; i < cnt; ++i) {
      double y, yd = 0.0, yg = 0.0;
      if (!coupled) {
        if (J->pathdep) sim_group<R, NF, 1, true>(J, 0, g, stage, tid);
        else sim_group<R, NF, 1, false>(J, 0, g, stage, tid);
        double pu = 0.0, pd = 0.0;
        eval3(J, stage, V_BASE, tid, K, GK, y, pu, pd);
        if constexpr (GK) {
          yd = (pu - pd) * J->inv2h;
          yg = (pu - y * 2.0 + pd) * J->invh2;
        }
      } else {
        sim_coupled<R, NF>(J, g, stage, tid);
        double pf, fu = 0.0, fd = 0.0, pc, cu = 0.0, cd = 0.0;
        eval3(J, stage, 0, tid, K, GK, pf, fu, fd);
        eval3(J, stage, 1, tid, K, GK, pc, cu, cd);
        y = pf - pc;
        if constexpr (GK) {
          yd = (fu - fd) * J->inv2h - (cu - cd) * J->inv2h;
          yg = (fu - pf * 2.0 + fd) * J->invh2 - (cu - pc * 2.0 + cd) * J->invh2;
        }
      }
      acc[0] += y;
      acc[1] += y * y;
      if constexpr (GK) {
        acc[2] += yd;
        acc[3] += yd * yd;
        acc[4] += yg;
        acc[5] += yg * yg;
      }
    }
    streams[4 * s] = g.s0;
    streams[4 * s + 1] = g.s1;
    streams[4 * s + 2] = g.s2;
    streams[4 * s + 3] = g.s3;
  }
  constexpr int NS = GK ? MLMC_SUMS : 2;
#pragma unroll
  for (int a = 0; a < NS; ++a)
    for (int off = 32; off > 0; off >>= 1) acc[a] += __shfl_xor(acc[a], off, 64);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < MLMC_SUMS; ++a) s_red[MLMC_SUMS * wave + a] = acc[a];
  }
  __syncthreads();
  if (tid < MLMC_SUMS) {
    double t = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) t += s_red[MLMC_SUMS * w + tid];
    block_sums[MLMC_SUMS * blockIdx.x + tid] = t;
  }
}

}  // namespace myql
