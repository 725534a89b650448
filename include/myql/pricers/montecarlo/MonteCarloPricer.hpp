// Drop-in MonteCarloPricer (reference include/myql/pricers/montecarlo/
// MonteCarloPricer.hpp:16-26 MonteCarloConfig, :84-488 the class). Same
// template parameters, constructor and calculate(S0, r, q, instrument)
// signature; calculate runs the whole pricing -- paths, payoffs, CRN Greek
// variants and the per-strike statistics -- in the HIP path kernels of
// libmyql_hip.so through myql_mc_calculate. Link with -lmyql_hip.
//
// As in the reference the pricer keeps a reference to the model (which must
// outlive it), copies the config, and returns the result by value. Errors
// raise std::runtime_error (SchemeNV with sigma^2 > 4 kappa theta, device
// failures) or std::invalid_argument.
#pragma once
#include <myql_capi.h>

#include <cstddef>
#include <type_traits>

#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Bridge.hpp>
#include <myql/detail/Results.hpp>
#include <myql/models/asvj/core/ASVJstepper.hpp>

struct MonteCarloConfig {
  size_t num_paths = 100000;
  size_t time_steps = 100;
  unsigned long seed = 42;
  double fd_bump = 1e-4;   // relative spot bump: h = S0 * fd_bump
  double vol_bump = 1e-4;  // absolute bump of v0 (vega, vanna)
  double t_bump = 1e-4;    // maturity bump (theta, charm)
  double r_bump = 1e-4;    // rate bump (rho)
  // Engine options (no counterpart in the reference; defaults give FP64
  // paths and the engine's default stream layout).
  int precision = MYQL_FP64;              // MYQL_FP64 / MYQL_FP32 path arithmetic
  int stream_layout = MYQL_STREAMS_JUMP;  // MYQL_STREAMS_JUMP / SEEDED / SOBOL (randomized QMC)
  size_t num_streams = 0;                 // 0: library default
  unsigned flags = 0;                     // MYQL_FLAG_*
  unsigned qmc_replicates = 0;            // SOBOL: independent digital shifts (0: 16)
};

namespace myql::detail {
inline myql_mc_config to_capi(const MonteCarloConfig &c) {
  myql_mc_config o{};
  o.num_paths = c.num_paths;
  o.time_steps = c.time_steps;
  o.seed = c.seed;
  o.fd_bump = c.fd_bump;
  o.vol_bump = c.vol_bump;
  o.t_bump = c.t_bump;
  o.r_bump = c.r_bump;
  o.precision = c.precision;
  o.stream_layout = c.stream_layout;
  o.num_streams = c.num_streams;
  o.flags = c.flags;
  o.qmc_replicates = c.qmc_replicates;
  return o;
}
}  // namespace myql::detail

// RNG is accepted for source compatibility only (the reference's default is
// boost::random::xoshiro256pp, which callers may spell out): whatever it
// names, the GPU engine draws from per-lane xoshiro256++ streams.
template <typename Model, typename Stepper, typename Instrument, GreekMode Mode = GreekMode::None,
          typename RNG = myql::xoshiro256pp>
class MonteCarloPricer {
  static_assert(std::is_same_v<typename Stepper::Policy, typename Instrument::Tracker>,
                "the stepper's tracker does not match the instrument's");

 public:
  using ResultType = typename Instrument::ResultType;
  using ReturnStruct = MonteCarloResult<Mode, ResultType>;

  MonteCarloPricer(const Model &model, const MonteCarloConfig &cfg) : model_(model), cfg_(cfg) {}

  ReturnStruct calculate(double S0, double r, double q, const Instrument &instr) const {
    int32_t schemes[2];
    const myql_model m = myql::detail::model_to_capi<Model, Stepper>(model_, schemes);
    const myql_instrument in = instr.to_capi();
    const myql_mc_config c = myql::detail::to_capi(cfg_);
    myql::detail::ResultArrays buf(instr.size());
    myql_mc_result res = buf.capi();
    myql_ctx *ctx = myql::detail::thread_context();
    myql::detail::check(
        myql_mc_calculate(ctx, &m, schemes, &in, myql::detail::greek_mode_id<Mode>, &c, S0, r, q, &res), ctx);
    ReturnStruct out{};
    buf.fill(out);
    out.total_steps = static_cast<size_t>(res.total_steps);
    return out;
  }

 private:
  const Model &model_;
  MonteCarloConfig cfg_;
};
