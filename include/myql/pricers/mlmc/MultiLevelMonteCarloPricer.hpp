// Drop-in MultiLevelMonteCarloPricer (reference include/myql/pricers/mlmc/
// MultiLevelMonteCarloPricer.hpp:16-30 MLMCConfig, :35-345 the class). The
// Giles level loop runs on the host inside myql_mlmc_calculate; every
// level's coupled fine/coarse paths run in the HIP level kernels. Scalar
// instruments only, as in the reference (the level differences are
// scalars). Under GreekMode::Essential and Full, delta and gamma are filled
// with their multilevel estimates (the reference leaves every Greek member
// zero); Full mode's other Greek members stay zero. MLMCConfig::estimator
// selects the multilevel Richardson-Romberg estimator (no counterpart in the
// reference). Link with -lmyql_hip.
#pragma once
#include <myql_capi.h>

#include <cstddef>
#include <type_traits>
#include <vector>

#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Bridge.hpp>
#include <myql/detail/Results.hpp>
#include <myql/models/asvj/core/CoupledASVJStepper.hpp>

struct MLMCConfig {
  double epsilon = 0.005;  // target RMSE
  size_t L_min = 2;
  size_t L_max = 15;
  size_t N0 = 10000;       // pilot paths at level 0
  size_t base_steps = 1;   // steps at level 0
  double alpha = 2.0;      // weak order
  double beta = 1.0;       // strong order
  bool adaptive_rates = true;
  unsigned long seed = 42;
  double fd_bump = 1e-4;
  // Engine options (no counterpart in the reference).
  int precision = MYQL_FP64;
  int stream_layout = MYQL_STREAMS_JUMP;
  size_t num_streams = 0;
  unsigned flags = 0;
  int estimator = MYQL_MLMC_GILES;  // or MYQL_MLMC_ML2R
};

template <typename Model, typename CoupledStepperType, typename Instrument,
          GreekMode Mode = GreekMode::None, typename RNG = myql::xoshiro256pp>
class MultiLevelMonteCarloPricer {
  static_assert(std::is_same_v<RNG, myql::xoshiro256pp>,
                "the GPU engine draws from per-lane xoshiro256++ streams only");
  static_assert(std::is_same_v<typename CoupledStepperType::Policy, typename Instrument::Tracker>,
                "the stepper's tracker does not match the instrument's");
  static_assert(std::is_floating_point_v<typename Instrument::ResultType>,
                "MLMC prices one strike at a time");

 public:
  using ResultType = typename Instrument::ResultType;
  using ReturnStruct = MlmcResult<Mode, ResultType>;

  MultiLevelMonteCarloPricer(const Model &model, const MLMCConfig &cfg) : model_(model), cfg_(cfg) {}

  ReturnStruct calculate(double S0, double r, double q, const Instrument &instr) const {
    int32_t schemes[2];
    const myql_model m = myql::detail::model_to_capi<Model, CoupledStepperType>(model_, schemes);
    const myql_instrument in = instr.to_capi();
    myql_mlmc_config c{};
    c.epsilon = cfg_.epsilon;
    c.L_min = cfg_.L_min;
    c.L_max = cfg_.L_max;
    c.N0 = cfg_.N0;
    c.base_steps = cfg_.base_steps;
    c.alpha = cfg_.alpha;
    c.beta = cfg_.beta;
    c.adaptive_rates = cfg_.adaptive_rates ? 1 : 0;
    c.seed = cfg_.seed;
    c.fd_bump = cfg_.fd_bump;
    c.precision = cfg_.precision;
    c.stream_layout = cfg_.stream_layout;
    c.num_streams = cfg_.num_streams;
    c.flags = cfg_.flags;
    c.estimator = cfg_.estimator;
    std::vector<uint64_t> paths(cfg_.L_max + 1, 0);
    std::vector<double> means(cfg_.L_max + 1, 0.0), vars(cfg_.L_max + 1, 0.0);
    myql_mlmc_result res{};
    res.paths_per_level = paths.data();
    res.mean_diffs = means.data();
    res.var_diffs = vars.data();
    myql_ctx *ctx = myql::detail::thread_context();
    myql::detail::check(myql_mlmc_calculate(ctx, &m, schemes, &in, myql::detail::greek_mode_id<Mode>, &c,
                                            S0, r, q, &res),
                        ctx);
    ReturnStruct out{};
    out.price = res.price;
    out.price_std_err = res.price_std_err;
    out.total_steps = static_cast<size_t>(res.total_steps);
    out.alpha_estimated = res.alpha_estimated;
    out.beta_estimated = res.beta_estimated;
    if constexpr (Mode != GreekMode::None) {
      out.delta = res.delta;
      out.delta_std_err = res.delta_std_err;
      out.gamma = res.gamma;
      out.gamma_std_err = res.gamma_std_err;
    }
    for (int l = 0; l < res.n_levels; ++l) {
      out.paths_per_level.push_back(static_cast<size_t>(paths[l]));
      out.mean_diffs.push_back(means[l]);
      out.var_diffs.push_back(vars[l]);
    }
    return out;
  }

 private:
  const Model &model_;
  MLMCConfig cfg_;
};
