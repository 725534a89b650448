// Instantiations and launch helpers of the learned control-variate kernels.
#include "cv_kernel.hpp"
#include "cv_launch.hpp"

namespace myql {

template <typename R>
hipError_t launch_cv_paths(int nf, const Job<R> *jobs, const CvSet *sets, const CvPathTask *tasks, int n_tasks,
                           unsigned blocks, hipStream_t stream) {
  const dim3 grid(blocks);
  switch (nf) {
    case 0: hipLaunchKernelGGL((cv_paths_kernel<R, 0>), grid, dim3(BLOCK), 0, stream, jobs, sets, tasks, n_tasks); break;
    case 1: hipLaunchKernelGGL((cv_paths_kernel<R, 1>), grid, dim3(BLOCK), 0, stream, jobs, sets, tasks, n_tasks); break;
    case 2: hipLaunchKernelGGL((cv_paths_kernel<R, 2>), grid, dim3(BLOCK), 0, stream, jobs, sets, tasks, n_tasks); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t launch_cv_paths<float>(int, const Job<float> *, const CvSet *, const CvPathTask *, int, unsigned,
                                           hipStream_t);
template hipError_t launch_cv_paths<double>(int, const Job<double> *, const CvSet *, const CvPathTask *, int,
                                            unsigned, hipStream_t);

template <typename R>
hipError_t launch_nci_rowmean(const R *tab, int rows, int grid, double dx, double d, double *mean,
                              hipStream_t stream) {
  hipLaunchKernelGGL(nci_rowmean_kernel<R>, dim3(rows), dim3(256), 0, stream, tab, rows, grid, dx, d, mean);
  return hipGetLastError();
}
template hipError_t launch_nci_rowmean<float>(const float *, int, int, double, double, double *, hipStream_t);
template hipError_t launch_nci_rowmean<double>(const double *, int, int, double, double, double *, hipStream_t);

hipError_t launch_cv_rb_table(const CvSet &set, double *tab, hipStream_t stream) {
  const int n = set.nk * CV_RB_PTS;
  hipLaunchKernelGGL(cv_rb_table_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, set, tab);
  return hipGetLastError();
}

hipError_t launch_cv_debug_rb(const CvSet &set, const double *spots, int64_t n, double *tab, double *dir,
                              double *ddir, hipStream_t stream) {
  const int64_t m = n * set.nk;
  hipLaunchKernelGGL(cv_debug_rb_kernel, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, stream, set, spots,
                     n, tab, dir, ddir);
  return hipGetLastError();
}

hipError_t launch_cv_moments(const CvSet &set, double *part, int blocks, hipStream_t stream) {
  hipLaunchKernelGGL(cv_moments_kernel, dim3(blocks, set.nk), dim3(256), 0, stream, set, part);
  return hipGetLastError();
}

template <int NF>
static void eval_nf(int mode, dim3 g, hipStream_t st, const float *prm, const CvNorm &nm, const CvSet *sets,
                    const CvEvalTask *tasks, int n) {
  if (mode == 0) hipLaunchKernelGGL((cv_eval_kernel<NF, 0>), g, dim3(BLOCK), 0, st, prm, nm, sets, tasks, n);
  else if (mode == 1) hipLaunchKernelGGL((cv_eval_kernel<NF, 1>), g, dim3(BLOCK), 0, st, prm, nm, sets, tasks, n);
  else hipLaunchKernelGGL((cv_eval_kernel<NF, 2>), g, dim3(BLOCK), 0, st, prm, nm, sets, tasks, n);
}

hipError_t launch_cv_eval(int nf, int mode, const float *prm, const CvNorm &nm, const CvSet *sets,
                          const CvEvalTask *tasks, int n_tasks, unsigned blocks, int max_nk, hipStream_t stream) {
  const dim3 g(blocks);
  if (nf <= 1) eval_nf<1>(mode, g, stream, prm, nm, sets, tasks, n_tasks);
  else eval_nf<2>(mode, g, stream, prm, nm, sets, tasks, n_tasks);
  const int ns = mode == 2 ? CV_NS_FIT : (mode == 1 ? CV_NS_PRICE : 4);
  hipLaunchKernelGGL(cv_reduce_kernel, dim3(max_nk * ns, n_tasks), dim3(256), 0, stream, sets, tasks,
                     ns);
  return hipGetLastError();
}

hipError_t launch_cv_train_step(const CvTrainArgs &a, hipStream_t stream) {
  const int nb = a.batch / CV_TROWS;
  hipLaunchKernelGGL(cv_train_kernel, dim3(nb), dim3(BLOCK), 0, stream, a.prm, a.sets, a.n_sets, a.norm, a.seed,
                     a.step, a.batch, a.nk_all, a.gpart, a.lpart, a.fixed_f, a.fixed_t);
  if (a.adam) {
    hipLaunchKernelGGL(cv_adam_kernel, dim3((CV_NPARAM + 255) / 256), dim3(256), 0, stream, a.prm, a.m, a.v,
                       a.gpart, nb, a.step, a.lr, a.lpart, a.loss_hist, a.loss_scale);
    hipLaunchKernelGGL(cv_step_kernel, dim3(1), dim3(64), 0, stream, a.step);
  }
  return hipGetLastError();
}

hipError_t launch_cv_grad_debug(const float *prm, const float *feats, int n, float *out, hipStream_t stream) {
  const int waves = (n + 15) / 16;
  const int blocks = (waves + BLOCK / 64 - 1) / (BLOCK / 64);
  hipLaunchKernelGGL(cv_grad_debug_kernel, dim3(blocks), dim3(BLOCK), 0, stream, prm, feats, n, out);
  return hipGetLastError();
}

}  // namespace myql
