// Instantiations of the path kernels with double path arithmetic.
#include "mc_kernel.hpp"
#include "mc_launch.hpp"

namespace myql {

#define MYQL_MC_CASES(X)                                                                      \
  X(0, 1, false) X(0, 5, false) X(1, 1, false) X(1, 5, false) X(2, 1, false) X(2, 7, false) \
  X(0, 1, true) X(0, 7, true) X(1, 1, true) X(1, 7, true) X(2, 1, true) X(2, 9, true)

template <>
hipError_t launch_mc<double>(int nf, int nv, bool pd, const Job<double> *jobs, const Unit *units, int n_units,
                           unsigned *counter, double *accg, size_t smem, int grid, hipStream_t stream) {
#define MYQL_LAUNCH(NF_, NV_, PD_)                                                                \
  if (nf == NF_ && nv == NV_ && pd == PD_) {                                                       \
    if (smem > 64 * 1024)                                                                          \
      (void)hipFuncSetAttribute(reinterpret_cast<const void *>(mc_paths_kernel<double, NF_, NV_, PD_>), \
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(smem)); \
    hipLaunchKernelGGL((mc_paths_kernel<double, NF_, NV_, PD_>), dim3(grid), dim3(BLOCK), smem, stream, \
                       jobs, units, n_units, counter, accg);                                      \
    return hipGetLastError();                                                                     \
  }
  MYQL_MC_CASES(MYQL_LAUNCH)
#undef MYQL_LAUNCH
  return hipErrorInvalidValue;
}

template <>
int mc_blocks_per_cu<double>(int nf, int nv, bool pd, size_t smem) {
  int nb = 0;
#define MYQL_OCC(NF_, NV_, PD_)                                                                      \
  if (nf == NF_ && nv == NV_ && pd == PD_) {                                                          \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mc_paths_kernel<double, NF_, NV_, PD_>, BLOCK, \
                                                     smem) != hipSuccess)                             \
      nb = 1;                                                                                         \
    return nb > 0 ? nb : 1;                                                                           \
  }
  MYQL_MC_CASES(MYQL_OCC)
#undef MYQL_OCC
  return 1;
}

}  // namespace myql
