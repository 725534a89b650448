// Drop-in runtime dispatch (reference include/myql/dispatcher/
// PricingDispatch.hpp:18-158): DispatchResult, price_mc and price_fourier
// over the model and instrument variants. Same result fields and function
// signatures as the reference.
//
// The reference resolves a (model, instrument) pair by a nested std::visit
// that instantiates a typed pricer for every pair. The GPU engine takes the
// model, variance schemes, jump law and tracker as run-time descriptors, so
// here the visits only read the descriptors out of the variants and one
// call prices any pair: the same C-ABI call the typed pricers make, so the
// numbers are theirs. time_ms, which the reference declares but never sets,
// holds the call's wall-clock time in milliseconds (descriptors, the GPU
// pricing and the copy of the results).
#pragma once
#include <myql_capi.h>

#include <chrono>
#include <variant>
#include <vector>

#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Bridge.hpp>
#include <myql/detail/Results.hpp>
#include <myql/dispatcher/InstrumentRegistry.hpp>
#include <myql/dispatcher/ModelRegistry.hpp>
#include <myql/dispatcher/StepperTraits.hpp>
#include <myql/engines/fourier/FourierEngine.hpp>
#include <myql/pricers/montecarlo/MonteCarloPricer.hpp>

namespace myql::dispatcher {

struct DispatchResult {
  std::vector<double> prices;
  // GreekMode::Essential and Full
  std::vector<double> deltas;
  std::vector<double> gammas;
  // GreekMode::Full: vegas[f] per variance factor f
  std::vector<std::vector<double>> vegas;
  std::vector<double> thetas;
  std::vector<double> rhos;
  // Monte Carlo standard errors (empty for Fourier)
  std::vector<double> prices_std_err;
  std::vector<double> deltas_std_err;
  std::vector<double> gammas_std_err;
  std::vector<std::vector<double>> vegas_std_err;
  std::vector<double> thetas_std_err;
  std::vector<double> rhos_std_err;

  double time_ms = 0.0;
};

namespace detail {

using Arrays = myql::detail::ResultArrays;

// The fields of `Mode` from the engine's per-strike arrays (scalar
// instruments give one-element vectors, as the reference wraps them);
// with_se: Monte Carlo standard errors too.
template <GreekMode Mode>
DispatchResult from_arrays(const Arrays &a, bool with_se) {
  using V = std::vector<double>;
  DispatchResult out;
  out.prices = a.get<V>(Arrays::PRICE);
  if (with_se) out.prices_std_err = a.get<V>(Arrays::PRICE_SE);
  if constexpr (Mode != GreekMode::None) {
    out.deltas = a.get<V>(Arrays::DELTA);
    out.gammas = a.get<V>(Arrays::GAMMA);
    if (with_se) {
      out.deltas_std_err = a.get<V>(Arrays::DELTA_SE);
      out.gammas_std_err = a.get<V>(Arrays::GAMMA_SE);
    }
  }
  if constexpr (Mode == GreekMode::Full) {
    out.vegas = {a.get<V>(Arrays::VEGA0), a.get<V>(Arrays::VEGA1)};
    out.thetas = a.get<V>(Arrays::THETA);
    out.rhos = a.get<V>(Arrays::RHO);
    if (with_se) {
      out.vegas_std_err = {a.get<V>(Arrays::VEGA_SE0), a.get<V>(Arrays::VEGA_SE1)};
      out.thetas_std_err = a.get<V>(Arrays::THETA_SE);
      out.rhos_std_err = a.get<V>(Arrays::RHO_SE);
    }
  }
  return out;
}

// Model descriptor; each variance factor takes the requested scheme (as the
// reference's StepperFor resolves it).
template <typename VolScheme>
myql_model model_desc(const AnyModel &m_var, int32_t schemes[2]) {
  return std::visit(
      [&](const auto &m) {
        constexpr int nf = std::decay_t<decltype(m)>::num_variance_factors;
        schemes[0] = nf >= 1 ? VolScheme::capi_id : NullVolScheme::capi_id;
        schemes[1] = nf == 2 ? VolScheme::capi_id : NullVolScheme::capi_id;
        return myql::detail::affine_model_to_capi(m);
      },
      m_var);
}

// Instrument descriptor and strike count. The descriptor points into the
// instrument, which the caller keeps alive through the call.
template <typename AnyI>
myql_instrument instrument_desc(const AnyI &i_var, size_t &n) {
  return std::visit(
      [&](const auto &i) {
        n = i.size();
        return i.to_capi();
      },
      i_var);
}

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace detail

template <GreekMode Mode = GreekMode::None, typename ExplicitVolScheme = SchemeNCI>
DispatchResult price_mc(const AnyModel &m_var, const AnyMCInstrument &i_var, const MonteCarloConfig &cfg, double S0,
                        double r, double q) {
  const auto t0 = std::chrono::steady_clock::now();
  int32_t schemes[2];
  const myql_model m = detail::model_desc<ExplicitVolScheme>(m_var, schemes);
  size_t n = 0;
  const myql_instrument in = detail::instrument_desc(i_var, n);
  const myql_mc_config c = myql::detail::to_capi(cfg);
  detail::Arrays buf(n);
  myql_mc_result res = buf.capi();
  myql_ctx *ctx = myql::detail::thread_context();
  myql::detail::check(
      myql_mc_calculate(ctx, &m, schemes, &in, myql::detail::greek_mode_id<Mode>, &c, S0, r, q, &res), ctx);
  DispatchResult out = detail::from_arrays<Mode>(buf, true);
  out.time_ms = detail::ms_since(t0);
  return out;
}

template <GreekMode Mode = GreekMode::None>
DispatchResult price_fourier(const AnyModel &m_var, const AnyEuropeanInstrument &i_var,
                             const FourierEngine::Config &cfg, double S0, double r, double q) {
  const auto t0 = std::chrono::steady_clock::now();
  const myql_model m = std::visit([](const auto &mm) { return myql::detail::affine_model_to_capi(mm); }, m_var);
  size_t n = 0;
  const myql_instrument in = detail::instrument_desc(i_var, n);
  const myql_fourier_config c{cfg.tolerance, cfg.start_bound_guess};
  detail::Arrays buf(n);
  myql_mc_result res = buf.capi();
  myql_fourier_job job{&m, &in, myql::detail::greek_mode_id<Mode>, &c, S0, r, q, &res};
  myql_ctx *ctx = myql::detail::thread_context();
  myql::detail::check(myql_fourier_calculate_batch(ctx, &job, 1), ctx);
  DispatchResult out = detail::from_arrays<Mode>(buf, false);
  out.time_ms = detail::ms_since(t0);
  return out;
}

}  // namespace myql::dispatcher
