// Geometric / arithmetic Asian option, fixed or floating strike, with the
// in-flight constructors (reference include/myql/instruments/options/
// Asian.hpp:7-131).
#pragma once
#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Strikes.hpp>
#include <myql/instruments/Payoffs.hpp>
#include <myql/instruments/trackers/PathTrackers.hpp>

template <typename TrackerT, typename PayoffT, bool FixedStrike, typename StrikeContainer = double>
class AsianOption {
  static_assert(std::is_same_v<TrackerT, TrackerGeoAsian> || std::is_same_v<TrackerT, TrackerArithAsian>,
                "AsianOption needs TrackerGeoAsian or TrackerArithAsian");

 public:
  using Tracker = TrackerT;
  using ResultType = StrikeContainer;
  using PayoffType = PayoffT;

  AsianOption(const StrikeContainer &K, double T, const PayoffT &payoff = PayoffT())
      : strikes_(K), T_(T), payoff_(payoff) {}
  // in flight (geometric): average of log S over the elapsed t_elapsed
  AsianOption(const StrikeContainer &K, double T, double past_avg_log, double t_elapsed,
              const PayoffT &payoff = PayoffT())
      : strikes_(K), T_(T), payoff_(payoff), past_avg_log_(past_avg_log), t_elapsed_(t_elapsed) {}
  // in flight (arithmetic): averages of S and of log S over t_elapsed
  AsianOption(const StrikeContainer &K, double T, double past_avg_S, double past_avg_log,
              double t_elapsed, const PayoffT &payoff = PayoffT())
      : strikes_(K), T_(T), payoff_(payoff), past_avg_log_(past_avg_log), past_avg_S_(past_avg_S),
        t_elapsed_(t_elapsed) {}

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
    o.past_avg_log = past_avg_log_;
    o.past_avg_S = past_avg_S_;
    o.t_elapsed = t_elapsed_;
    myql::detail::set_payoff(o, payoff_);
    myql::detail::attach_strikes(o, strikes_);
    return o;
  }

 private:
  StrikeContainer strikes_;
  double T_;
  PayoffT payoff_;
  double past_avg_log_ = 0.0;
  double past_avg_S_ = 0.0;
  double t_elapsed_ = 0.0;
};
