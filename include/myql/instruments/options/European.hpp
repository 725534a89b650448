// European option, one strike or a strip (reference include/myql/
// instruments/options/European.hpp:10-86).
#pragma once
#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Strikes.hpp>
#include <myql/instruments/Payoffs.hpp>
#include <myql/instruments/trackers/PathTrackers.hpp>

template <typename PayoffPolicy, typename StrikeContainer = double>
class EuropeanOption {
 public:
  using Tracker = TrackerEuropean;
  using ResultType = StrikeContainer;
  using PayoffType = PayoffPolicy;

  EuropeanOption(const StrikeContainer &K, double T, const PayoffPolicy &payoff = PayoffPolicy())
      : strikes_(K), T_(T), payoff_(payoff) {}

  double get_maturity() const { return T_; }
  const StrikeContainer &get_strikes() const { return strikes_; }
  PayoffPolicy &get_payoff_mut() { return payoff_; }
  size_t size() const {
    if constexpr (myql::detail::is_scalar_strike<StrikeContainer>) return 1;
    else return strikes_.size();
  }

  myql_instrument to_capi() const {
    myql_instrument o = myql::detail::blank_instrument(T_);
    o.kind = Tracker::capi_kind;
    myql::detail::set_payoff(o, payoff_);
    myql::detail::attach_strikes(o, strikes_);
    return o;
  }

 private:
  StrikeContainer strikes_;
  double T_;
  PayoffPolicy payoff_;
};
