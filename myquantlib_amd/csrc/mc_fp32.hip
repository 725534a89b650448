// Instantiations of the path kernels with float path arithmetic.
#include "mc_kernel.hpp"
#include "mc_launch.hpp"

namespace myql {

#define MYQL_MC_CASES(X) X(0, 1) X(0, 5) X(0, 7) X(1, 1) X(1, 5) X(1, 7) X(2, 1) X(2, 7) X(2, 9)

template <>
hipError_t launch_mc<float>(int nf, int nv, const Job<float> *jobs, const Unit *units, int n_units,
                           unsigned *counter, size_t smem, int grid, hipStream_t stream) {
#define MYQL_LAUNCH(NF_, NV_)                                                                  \
  if (nf == NF_ && nv == NV_) {                                                                \
    hipLaunchKernelGGL((mc_paths_kernel<float, NF_, NV_>), dim3(grid), dim3(BLOCK), smem, stream, \
                       jobs, units, n_units, counter);                                         \
    return hipGetLastError();                                                                  \
  }
  MYQL_MC_CASES(MYQL_LAUNCH)
#undef MYQL_LAUNCH
  return hipErrorInvalidValue;
}

template <>
int mc_blocks_per_cu<float>(int nf, int nv, size_t smem) {
  int nb = 0;
#define MYQL_OCC(NF_, NV_)                                                                         \
  if (nf == NF_ && nv == NV_) {                                                                    \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mc_paths_kernel<float, NF_, NV_>, BLOCK, \
                                                     smem) != hipSuccess)                          \
      nb = 1;                                                                                      \
    return nb > 0 ? nb : 1;                                                                        \
  }
  MYQL_MC_CASES(MYQL_OCC)
#undef MYQL_OCC
  return 1;
}

}  // namespace myql
