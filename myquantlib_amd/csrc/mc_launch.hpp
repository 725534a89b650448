// Launch helpers for the path kernels; one translation unit per path
// precision instantiates them (mc_fp32.hip, mc_fp64.hip via mc_inst.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace myql {

// Lockstep widths compiled for each (factor count, path dependence):
//   European NF 0/1: 1, 5; NF 2: 1, 7. Path-dependent NF 0/1: 1, 7; NF 2: 1, 9.
bool mc_supported(int nf, int nv, bool pd);

// Launch shape of the path kernel: streams per work unit (= lanes per block),
// the unroll of the strike loop over staged samples, and occ = 4 for the
// variant compiled to fit four waves per SIMD (0: the compiler's choice).
// Every kernel is compiled at three block sizes with the loop unrolled by 4
// (the first, 256/4, is the default shape); the lockstep European kernels
// (the strip workloads, whose strike loop dominates the staging phase) also
// with the loop not unrolled and, at 64-stream units, in the four-wave
// variant. The autotuner picks among them.
// qmc = 1: the kernel drawing randomized quasi-Monte Carlo points
// (mc_paths_kernel_qmc), compiled at the default shape only.
struct McShape {
  int blk, unr, occ;
  int qmc = 0;
};
constexpr McShape MC_DEFAULT_SHAPE{BLOCK, 4, 0};
constexpr McShape MC_QMC_SHAPE{BLOCK, 4, 0, 1};
constexpr McShape MC_TUNE_SHAPES[8] = {{256, 4, 0}, {128, 4, 0}, {64, 4, 0}, {256, 1, 0},
                                       {128, 1, 0}, {64, 1, 0},  {64, 4, 4}, {64, 1, 4}};
// number of leading MC_TUNE_SHAPES compiled for a kernel
inline int mc_tune_shape_count(int nf, int nv, bool pd) { return (!pd && nv > 1 && nf >= 0 && nf <= 2) ? 8 : 3; }

// pd: path-dependent tracker kernels; accg: 2*NSTAT*shape.blk doubles per unit.
template <typename R>
hipError_t launch_mc(int nf, int nv, bool pd, McShape shape, const Job<R> *jobs, const Unit *units, int n_units,
                     unsigned *counter, double *accg, size_t smem, int grid, hipStream_t stream);

template <typename R>
int mc_blocks_per_cu(int nf, int nv, bool pd, McShape shape, size_t smem);

}  // namespace myql
