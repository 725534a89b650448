// Greek modes and result structs of the MI355X engine's C++20 layer.
// Same names, template parameters and members as the reference's
// include/myql/core/PricingTypes.hpp:8-168 (GreekMode, MonteCarloResult,
// MlmcResult), so code written against the reference compiles unchanged.
// T is double for a single strike or std::vector<double> for a strip.
#pragma once
#include <array>
#include <cstddef>
#include <vector>

enum class GreekMode { None, Essential, Full };

// FourierResult (reference PricingTypes.hpp:17-41): price and, per mode,
// delta, gamma, vega[2], theta (dP/dT), rho, vanna[2], charm.
template <GreekMode Mode, typename T = double>
struct FourierResult;
template <typename T>
struct FourierResult<GreekMode::None, T> {
  T price{};
};
template <typename T>
struct FourierResult<GreekMode::Essential, T> {
  T price{};
  T delta{};
  T gamma{};
};
template <typename T>
struct FourierResult<GreekMode::Full, T> {
  T price{};
  T delta{};
  T gamma{};
  std::array<T, 2> vega{};
  T theta{};
  T rho{};
  std::array<T, 2> vanna{};
  T charm{};
};

template <GreekMode Mode, typename T>
struct MonteCarloResult;

template <typename T>
struct MonteCarloResult<GreekMode::None, T> {
  T price;
  T price_std_err;
  size_t total_steps{};
};

template <typename T>
struct MonteCarloResult<GreekMode::Essential, T> {
  T price;
  T price_std_err;
  T delta;
  T delta_std_err;
  T gamma;
  T gamma_std_err;
  size_t total_steps{};
};

template <typename T>
struct MonteCarloResult<GreekMode::Full, T> {
  T price;
  T price_std_err;
  T delta;
  T delta_std_err;
  T gamma;
  T gamma_std_err;
  std::array<T, 2> vega{};  // per variance factor; [1] is 0 for one factor
  std::array<T, 2> vega_std_err{};
  T theta;
  T theta_std_err;
  T rho;
  T rho_std_err;
  std::array<T, 2> vanna{};
  std::array<T, 2> vanna_std_err{};
  T charm;
  T charm_std_err;
  size_t total_steps{};
};

// MLMC results. As in the reference, only the price, its standard error
// and the level statistics are produced; the Greek members of the
// Essential / Full forms exist for source compatibility and stay zero.
#define MYQL_MLMC_LEVEL_MEMBERS          \
  size_t total_steps{};                  \
  std::vector<size_t> paths_per_level{}; \
  std::vector<T> mean_diffs{};           \
  std::vector<T> var_diffs{};            \
  double alpha_estimated{};              \
  double beta_estimated{};

template <GreekMode Mode, typename T>
struct MlmcResult;

template <typename T>
struct MlmcResult<GreekMode::None, T> {
  T price{};
  T price_std_err{};
  MYQL_MLMC_LEVEL_MEMBERS
};

template <typename T>
struct MlmcResult<GreekMode::Essential, T> {
  T price{};
  T price_std_err{};
  T delta{};
  T delta_std_err{};
  T gamma{};
  T gamma_std_err{};
  MYQL_MLMC_LEVEL_MEMBERS
};

template <typename T>
struct MlmcResult<GreekMode::Full, T> {
  T price{};
  T price_std_err{};
  T delta{};
  T delta_std_err{};
  T gamma{};
  T gamma_std_err{};
  std::array<T, 2> vega{};
  std::array<T, 2> vega_std_err{};
  T theta{};
  T theta_std_err{};
  T rho{};
  T rho_std_err{};
  std::array<T, 2> vanna{};
  std::array<T, 2> vanna_std_err{};
  T charm{};
  T charm_std_err{};
  MYQL_MLMC_LEVEL_MEMBERS
};

#undef MYQL_MLMC_LEVEL_MEMBERS
