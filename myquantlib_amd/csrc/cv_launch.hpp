// Launch helpers of the learned control-variate kernels (cv.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cv_types.hpp"
#include "engine_types.hpp"

namespace myql {

struct CvTrainArgs {
  float *prm, *m, *v, *gpart, *lpart, *loss_hist;
  const CvSet *sets;  // device array
  int n_sets;
  CvNorm norm;
  uint64_t seed;
  int *step;          // device counter
  int batch;          // rows per step, multiple of 32
  int nk_all;         // strikes per set (the same for every set)
  float lr;
  float loss_scale;   // 1 / blocks: the recorded loss is the mean over blocks
  bool adam;          // false: gradient only (checks)
  const float *fixed_f, *fixed_t;  // caller rows instead of sampled ones (or null)
};

// jobs, sets, tasks: device arrays; blocks: the grid (the last task's
// first_block plus its blocks)
template <typename R>
hipError_t launch_cv_paths(int nf, const Job<R> *jobs, const CvSet *sets, const CvPathTask *tasks, int n_tasks,
                           unsigned blocks, hipStream_t stream);
template <typename R>
hipError_t launch_nci_rowmean(const R *tab, int rows, int grid, double dx, double d, double *mean,
                              hipStream_t stream);
hipError_t launch_cv_rb_table(const CvSet &set, double *tab, hipStream_t stream);
hipError_t launch_cv_debug_rb(const CvSet &set, const double *spots, int64_t n, double *tab, double *dir,
                              double *ddir, hipStream_t stream);
hipError_t launch_cv_moments(const CvSet &set, double *part, int blocks, hipStream_t stream);
// mode 0: price only, 1: with plain delta and gamma, 2: fitting sums; each
// task's `sums` receives nk x ns sums (ns = 4, 8 or 21); tasks: device
// array; max_nk: the largest strike count of the tasks
hipError_t launch_cv_eval(int nf, int mode, const float *prm, const CvNorm &nm, const CvSet *sets,
                          const CvEvalTask *tasks, int n_tasks, unsigned blocks, int max_nk, hipStream_t stream);
hipError_t launch_cv_train_step(const CvTrainArgs &a, hipStream_t stream);
hipError_t launch_cv_grad_debug(const float *prm, const float *feats, int n, float *out, hipStream_t stream);

}  // namespace myql
