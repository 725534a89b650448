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
  float lr;
  bool adam;          // false: gradient only (checks)
  const float *fixed_f, *fixed_t;  // caller rows instead of sampled ones (or null)
};

template <typename R>
hipError_t launch_cv_paths(int nf, const Job<R> *job, const CvSet &set, uint64_t S, hipStream_t stream);
hipError_t launch_cv_eval(const float *prm, const CvNorm &nm, const CvSet &set, double *part, int blocks,
                          double *sums, hipStream_t stream);
hipError_t launch_cv_train_step(const CvTrainArgs &a, hipStream_t stream);
hipError_t launch_cv_grad_debug(const float *prm, const float *feats, int n, float *out, hipStream_t stream);

}  // namespace myql
