// Drop-in FourierEngine configuration (reference include/myql/engines/
// fourier/FourierEngine.hpp:8-18). The integration itself runs on the GPU
// (myql_fourier_calculate_batch); only the Config type is part of the API.
#pragma once

struct FourierConfig {
  double tolerance = 1e-8;         // the reference's adaptive-Simpson target (accepted, unused)
  double start_bound_guess = 20.0; // first point of the integration-bound search
};

class FourierEngine {
 public:
  using Config = FourierConfig;
};
