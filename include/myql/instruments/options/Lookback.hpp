// Lookback option, fixed or floating strike, optionally seeded with the
// observed extremes (reference include/myql/instruments/options/
// Lookback.hpp:8-111).
#pragma once
#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Strikes.hpp>
#include <myql/instruments/Payoffs.hpp>
#include <myql/instruments/trackers/PathTrackers.hpp>

template <typename PayoffT, bool FixedStrike, typename StrikeContainer = double>
class LookbackOption {
 public:
  using Tracker = TrackerLookback;
  using ResultType = StrikeContainer;
  using PayoffType = PayoffT;

  // hist_min_S / hist_max_S: observed extremes so far (0 = no history)
  LookbackOption(const StrikeContainer &K, double T, double hist_min_S = 0.0, double hist_max_S = 0.0,
                 const PayoffT &payoff = PayoffT())
      : strikes_(K), T_(T), payoff_(payoff), hist_min_S_(hist_min_S), hist_max_S_(hist_max_S) {}

  double get_maturity() const { return T_; }
  const StrikeContainer &get_strikes() const { return strikes_; }
  PayoffT &get_payoff_mut() { return payoff_; }
  size_t size() const {
    if constexpr (myql::detail::is_scalar_strike<StrikeContainer>) return 1;
    else return strikes_.size();
  }

  myql_instrument to_capi() const {
    myql_instrument o = myql::detail::blank_instrument(T_);
    o.kind = Tracker::capi_kind;
    o.fixed_strike = FixedStrike ? 1 : 0;
    o.hist_min_S = hist_min_S_;
    o.hist_max_S = hist_max_S_;
    myql::detail::set_payoff(o, payoff_);
    myql::detail::attach_strikes(o, strikes_);
    return o;
  }

 private:
  StrikeContainer strikes_;
  double T_;
  PayoffT payoff_;
  double hist_min_S_;
  double hist_max_S_;
};
