// Host-device interface of the Gil-Pelaez Fourier pricer (fourier.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace myql {

// ASVJ model as the characteristic exponent needs it (ASVJmodel.hpp:14-87):
// h[f] = {kappa, theta, sigma, rho, v0}; jump = Merton {lambda, mu, delta} or
// Kou {lambda, p_up, eta1, eta2}.
struct FourierModel {
  int nf, jump_kind;
  double vol;
  double h[2][5];
  double jump[4];
};

// One option (one strike of one job): drift = (r - q) T, logK = log(K / S0),
// start_bound = the bound search's first point (start_bound_guess / sqrt(T)
// for T < 1); mode 2 also integrates the vega, theta and cross kernels.
struct FourierOption {
  FourierModel model;
  double T, drift, logK, start_bound;
  int mode;
};

// Integrals per option, each divided by pi, laid out [P1 | P2] x {price, Dx,
// Dxx, vega f0, vega f1, theta, vega-Dx f0, vega-Dx f1, theta-Dx}.
int fourier_n_integrals();
hipError_t launch_fourier(const FourierOption *opts, int n_opts, double *out, hipStream_t stream);

}  // namespace myql
