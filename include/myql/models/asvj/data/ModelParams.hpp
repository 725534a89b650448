// Model parameter atoms (reference include/myql/models/asvj/data/
// ModelParams.hpp:7-47): Heston diffusion and the jump laws.
#pragma once

struct HestonParams {
  double kappa;  // mean reversion
  double theta;  // long-run variance
  double sigma;  // vol of variance
  double rho;    // spot/variance correlation
  double v0;     // initial variance
  // SchemeNV is defined only for sigma^2 <= 4 kappa theta
  bool is_low_vol_regime() const { return sigma * sigma <= 4.0 * kappa * theta; }
  bool is_Feller_regime() const { return sigma * sigma <= 2.0 * kappa * theta; }
};

struct NoJumpParams {};

struct MertonParams {
  double lambda;  // intensity
  double mu;      // mean log jump
  double delta;   // std dev of the log jump
};

struct KouParams {
  double lambda;  // intensity
  double p_up;    // probability of an up jump
  double eta1;    // rate of up jumps
  double eta2;    // rate of down jumps
};
