// Parameter blocks of the learned control-variate kernels (cv_kernel.hpp),
// shared with the host orchestration (engine.hip).
#pragma once
#include <stdint.h>

namespace myql {

constexpr int CV_H = 64;
constexpr int CV_IN = 6;
constexpr int CV_NPARAM = 8833;
constexpr int CV_W1 = 0, CV_B1 = 384, CV_W2 = 448, CV_B2 = 4544, CV_W3 = 4608, CV_B3 = 8704, CV_W4 = 8768,
              CV_B4 = 8832;
constexpr int CV_MAXJ = 16;       // control points per path
constexpr int CV_REC = 6;         // recorded fields per control point
constexpr int CV_PITCH = 68;      // bf16 row pitch of the LDS weight copies (bank spread)
constexpr int CV_TROWS = 32;      // rows per training block

// Feature scaling: f0 = t*ts, f1 = (x - log K)*xs, f2 = v1*vs0, f3 = v2*vs1,
// f4 = T*ts, f5 = (log K - log S0)*xs; u = S0 * net(f).
struct CvNorm {
  float ts, xs, vs0, vs1, logS0, S0;
};

// One option set (a maturity's strikes) of a control-variate run.
struct CvSet {
  // recorded paths: rec[(j*CV_REC + c)*M + p], c = x, v1, v2, dW_perp, dW_v1,
  // dW_v2 at control point j; ST[p] terminal spot
  float *rec;
  double *ST;
  float *dJ;                    // e^{jump total} - 1 of the terminal jump (mean 0), or null
  uint64_t M;
  int nj;                       // control points
  int step_of[CV_MAXJ + 1];     // first step of each block, step_of[nj] = steps
  float t_of[CV_MAXJ];          // time of each control point
  double disc_of[CV_MAXJ];      // e^{r (T - t_j)}
  float sqdt;
  // model coefficients of sigma(X): stock and variance diffusion
  int nf;
  float vol;                    // zero-factor volatility
  float rho[2], sig[2], rperp[2];  // rperp = 1 - rho^2
  // instrument
  const double *strikes;        // device
  int nk, payoff, call;
  double T, df;
  // training targets
  double inv_S0;
};

}  // namespace myql
