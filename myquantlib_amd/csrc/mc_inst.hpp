// Instantiations of the path kernels for one path precision, MYQL_R, which
// the including translation unit defines (mc_fp32.hip: float, mc_fp64.hip:
// double).
#include "mc_kernel.hpp"
#include "mc_launch.hpp"

namespace myql {

// (factor count, lockstep width, path dependence, block, unroll): every
// kernel at blocks 256, 128, 64 with unroll 4, the lockstep European kernels
// also with unroll 1 (MC_TUNE_SHAPES).
#define MYQL_MC_KERNELS(X, B_, U_)                                                                      \
  X(0, 1, false, B_, U_) X(0, 5, false, B_, U_) X(1, 1, false, B_, U_) X(1, 5, false, B_, U_)           \
  X(2, 1, false, B_, U_) X(2, 7, false, B_, U_) X(0, 1, true, B_, U_) X(0, 7, true, B_, U_)             \
  X(1, 1, true, B_, U_) X(1, 7, true, B_, U_) X(2, 1, true, B_, U_) X(2, 9, true, B_, U_)
#define MYQL_MC_LOCKSTEP_EURO(X, B_, U_) X(0, 5, false, B_, U_) X(1, 5, false, B_, U_) X(2, 7, false, B_, U_)
#define MYQL_MC_CASES(X)                                                                                  \
  MYQL_MC_KERNELS(X, 256, 4) MYQL_MC_KERNELS(X, 128, 4) MYQL_MC_KERNELS(X, 64, 4)                       \
  MYQL_MC_LOCKSTEP_EURO(X, 256, 1) MYQL_MC_LOCKSTEP_EURO(X, 128, 1) MYQL_MC_LOCKSTEP_EURO(X, 64, 1)
// four-wave variants (occ 4)
#define MYQL_MC_CASES_OCC4(X) MYQL_MC_LOCKSTEP_EURO(X, 64, 4) MYQL_MC_LOCKSTEP_EURO(X, 64, 1)

template <>
hipError_t launch_mc<MYQL_R>(int nf, int nv, bool pd, McShape shape, const Job<MYQL_R> *jobs, const Unit *units,
                             int n_units, unsigned *counter, double *accg, size_t smem, int grid,
                             hipStream_t stream) {
#define MYQL_LAUNCH_Q(NF_, NV_, PD_, B_, U_)                                                                  \
  if (nf == NF_ && nv == NV_ && pd == PD_) {                                                                 \
    hipLaunchKernelGGL((mc_paths_kernel_qmc<MYQL_R, NF_, NV_, PD_>), dim3(grid), dim3(BLOCK), smem, stream,   \
                       jobs, units, n_units, counter, accg);                                                 \
    return hipGetLastError();                                                                                \
  }
  if (shape.qmc) {
    MYQL_MC_KERNELS(MYQL_LAUNCH_Q, BLOCK, 4)
    return hipErrorInvalidValue;
  }
#undef MYQL_LAUNCH_Q
#define MYQL_LAUNCH_K(KERNEL, OCC_, NF_, NV_, PD_, B_, U_)                                                   \
  if (!shape.qmc && nf == NF_ && nv == NV_ && pd == PD_ && shape.blk == B_ && shape.unr == U_ &&              \
      shape.occ == OCC_) {                                                                                   \
    if (smem > 64 * 1024)                                                                                    \
      (void)hipFuncSetAttribute(reinterpret_cast<const void *>(KERNEL<MYQL_R, NF_, NV_, PD_, B_, U_>),        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(smem));         \
    hipLaunchKernelGGL((KERNEL<MYQL_R, NF_, NV_, PD_, B_, U_>), dim3(grid), dim3(B_), smem, stream, jobs,     \
                       units, n_units, counter, accg);                                                       \
    return hipGetLastError();                                                                                \
  }
#define MYQL_LAUNCH(NF_, NV_, PD_, B_, U_) MYQL_LAUNCH_K(mc_paths_kernel, 0, NF_, NV_, PD_, B_, U_)
#define MYQL_LAUNCH_OCC4(NF_, NV_, PD_, B_, U_) MYQL_LAUNCH_K(mc_paths_kernel_occ4, 4, NF_, NV_, PD_, B_, U_)
  MYQL_MC_CASES(MYQL_LAUNCH)
  MYQL_MC_CASES_OCC4(MYQL_LAUNCH_OCC4)
#undef MYQL_LAUNCH
#undef MYQL_LAUNCH_OCC4
#undef MYQL_LAUNCH_K
  return hipErrorInvalidValue;
}

template <>
int mc_blocks_per_cu<MYQL_R>(int nf, int nv, bool pd, McShape shape, size_t smem) {
  int nb = 0;
#define MYQL_OCC_Q(NF_, NV_, PD_, B_, U_)                                                                     \
  if (nf == NF_ && nv == NV_ && pd == PD_) {                                                                  \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mc_paths_kernel_qmc<MYQL_R, NF_, NV_, PD_>, BLOCK,   \
                                                     smem) != hipSuccess)                                     \
      nb = 1;                                                                                                 \
    return nb > 0 ? nb : 1;                                                                                   \
  }
  if (shape.qmc) {
    MYQL_MC_KERNELS(MYQL_OCC_Q, BLOCK, 4)
    return 1;
  }
#undef MYQL_OCC_Q
#define MYQL_OCC_K(KERNEL, OCC_, NF_, NV_, PD_, B_, U_)                                                       \
  if (!shape.qmc && nf == NF_ && nv == NV_ && pd == PD_ && shape.blk == B_ && shape.unr == U_ &&               \
      shape.occ == OCC_) {                                                                                    \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, KERNEL<MYQL_R, NF_, NV_, PD_, B_, U_>, B_, smem) !=   \
        hipSuccess)                                                                                            \
      nb = 1;                                                                                                  \
    return nb > 0 ? nb : 1;                                                                                    \
  }
#define MYQL_OCC(NF_, NV_, PD_, B_, U_) MYQL_OCC_K(mc_paths_kernel, 0, NF_, NV_, PD_, B_, U_)
#define MYQL_OCC4(NF_, NV_, PD_, B_, U_) MYQL_OCC_K(mc_paths_kernel_occ4, 4, NF_, NV_, PD_, B_, U_)
  MYQL_MC_CASES(MYQL_OCC)
  MYQL_MC_CASES_OCC4(MYQL_OCC4)
#undef MYQL_OCC
#undef MYQL_OCC4
#undef MYQL_OCC_K
  return 1;
}

#undef MYQL_MC_CASES
#undef MYQL_MC_CASES_OCC4
#undef MYQL_MC_KERNELS
#undef MYQL_MC_LOCKSTEP_EURO

}  // namespace myql
