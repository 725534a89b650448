// Instantiations and launch helpers of the learned control-variate kernels.
#include "cv_kernel.hpp"
#include "cv_launch.hpp"

namespace myql {

template <typename R>
hipError_t launch_cv_paths(int nf, const Job<R> *job, const CvSet &set, uint64_t S, hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>((S + BLOCK - 1) / BLOCK));
  switch (nf) {
    case 0: hipLaunchKernelGGL((cv_paths_kernel<R, 0>), grid, dim3(BLOCK), 0, stream, job, set); break;
    case 1: hipLaunchKernelGGL((cv_paths_kernel<R, 1>), grid, dim3(BLOCK), 0, stream, job, set); break;
    case 2: hipLaunchKernelGGL((cv_paths_kernel<R, 2>), grid, dim3(BLOCK), 0, stream, job, set); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t launch_cv_paths<float>(int, const Job<float> *, const CvSet &, uint64_t, hipStream_t);
template hipError_t launch_cv_paths<double>(int, const Job<double> *, const CvSet &, uint64_t, hipStream_t);

hipError_t launch_cv_eval(const float *prm, const CvNorm &nm, const CvSet &set, double *part, int blocks,
                          double *sums, hipStream_t stream) {
  hipLaunchKernelGGL(cv_eval_kernel, dim3(blocks), dim3(BLOCK), 0, stream, prm, nm, set, part);
  const int n = set.nk * 4;
  hipLaunchKernelGGL(cv_reduce_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, part, blocks * (BLOCK / 64),
                     set.nk, sums);
  return hipGetLastError();
}

hipError_t launch_cv_train_step(const CvTrainArgs &a, hipStream_t stream) {
  const int nb = a.batch / CV_TROWS;
  hipLaunchKernelGGL(cv_train_kernel, dim3(nb), dim3(BLOCK), 0, stream, a.prm, a.sets, a.n_sets, a.norm, a.seed,
                     a.step, a.batch, a.gpart, a.lpart, a.fixed_f, a.fixed_t);
  if (a.adam) {
    hipLaunchKernelGGL(cv_adam_kernel, dim3((CV_NPARAM + 255) / 256), dim3(256), 0, stream, a.prm, a.m, a.v,
                       a.gpart, nb, a.step, a.lr, a.lpart, a.loss_hist, a.batch);
    hipLaunchKernelGGL(cv_step_kernel, dim3(1), dim3(64), 0, stream, a.step);
  }
  return hipGetLastError();
}

hipError_t launch_cv_grad_debug(const float *prm, const float *feats, int n, float *out, hipStream_t stream) {
  const int waves = (n + 31) / 32;
  const int blocks = (waves + BLOCK / 64 - 1) / (BLOCK / 64);
  hipLaunchKernelGGL(cv_grad_debug_kernel, dim3(blocks), dim3(BLOCK), 0, stream, prm, feats, n, out);
  return hipGetLastError();
}

}  // namespace myql
