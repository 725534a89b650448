// Drop-in FourierPricer (reference include/myql/pricers/fourier/
// FourierPricer.hpp:34-407): same template parameters, constructor and
// calculate(S0, r, q, instrument) for European vanilla, cash-or-nothing and
// asset-or-nothing options, scalar or strike strip. The Gil-Pelaez integrals
// of every strike run in one GPU launch (myql_fourier_calculate_batch);
// results are FourierResult<Mode, T> as in the reference. Link with
// -lmyql_hip. The Traits parameter is accepted for source compatibility; the
// characteristic exponents are fixed by the model type.
#pragma once
#include <myql_capi.h>

#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Bridge.hpp>
#include <myql/detail/Results.hpp>
#include <myql/engines/fourier/FourierEngine.hpp>
#include <myql/instruments/options/European.hpp>
#include <myql/models/asvj/core/ASVJmodel.hpp>

template <typename Model, typename Instrument, GreekMode Mode = GreekMode::None, typename Traits = void>
class FourierPricer {
 public:
  using ResultType = typename Instrument::ResultType;
  using ReturnStruct = FourierResult<Mode, ResultType>;

  FourierPricer(const Model &model, const FourierEngine::Config &cfg = FourierEngine::Config{})
      : model_(model), cfg_(cfg) {}

  ReturnStruct calculate(double S0, double r, double q, const Instrument &instr) const {
    const myql_model m = myql::detail::affine_model_to_capi(model_);
    const myql_instrument in = instr.to_capi();
    const myql_fourier_config c{cfg_.tolerance, cfg_.start_bound_guess};
    myql::detail::ResultArrays buf(instr.size());
    myql_mc_result res = buf.capi();
    myql_fourier_job job{&m, &in, myql::detail::greek_mode_id<Mode>, &c, S0, r, q, &res};
    myql_ctx *ctx = myql::detail::thread_context();
    myql::detail::check(myql_fourier_calculate_batch(ctx, &job, 1), ctx);
    ReturnStruct out{};
    buf.fill(out);
    return out;
  }

 private:
  const Model &model_;
  FourierEngine::Config cfg_;
};
