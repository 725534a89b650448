// Launch helpers for the path kernels; one translation unit per path
// precision instantiates them (mc_fp32.hip, mc_fp64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace myql {

// Lockstep widths compiled for each (factor count, path dependence):
//   European NF 0/1: 1, 5; NF 2: 1, 7. Path-dependent NF 0/1: 1, 7; NF 2: 1, 9.
bool mc_supported(int nf, int nv, bool pd);

// pd: path-dependent tracker kernels; accg: 2*NSTAT*BLOCK doubles per unit.
template <typename R>
hipError_t launch_mc(int nf, int nv, bool pd, const Job<R> *jobs, const Unit *units, int n_units,
                     unsigned *counter, double *accg, size_t smem, int grid, hipStream_t stream);

template <typename R>
int mc_blocks_per_cu(int nf, int nv, bool pd, size_t smem);

}  // namespace myql
