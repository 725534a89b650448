// Single barrier option with optional rebate, discretely monitored at the
// time steps (reference include/myql/instruments/options/Barrier.hpp:10-170).
// `r` is the rate used for an at-hit rebate's accrual (the instrument's own
// rate, not the pricing call's).
#pragma once
#include <myql/core/PricingTypes.hpp>
#include <myql/detail/Strikes.hpp>
#include <myql/instruments/Payoffs.hpp>
#include <myql/instruments/trackers/PathTrackers.hpp>

enum class BarrierDirection { Up, Down };
enum class BarrierAction { In, Out };
enum class RebateTiming { None, AtHit, AtMaturity };

template <typename PayoffT, BarrierDirection Dir, BarrierAction Action, RebateTiming Timing,
          typename StrikeContainer = double>
class BarrierOption {
 public:
  using Tracker = TrackerBarrier<(Dir == BarrierDirection::Up)>;
  using ResultType = StrikeContainer;
  using PayoffType = PayoffT;

  BarrierOption(const StrikeContainer &K, double B, double T, double r = 0.0, double rebate = 0.0,
                const PayoffT &payoff = PayoffT())
      : strikes_(K), barrier_(B), T_(T), rate_(r), rebate_(rebate), payoff_(payoff) {}

  double get_maturity() const { return T_; }
  double get_barrier() const { return barrier_; }
  const StrikeContainer &get_strikes() const { return strikes_; }
  PayoffT &get_payoff_mut() { return payoff_; }
  size_t size() const {
    if constexpr (myql::detail::is_scalar_strike<StrikeContainer>) return 1;
    else return strikes_.size();
  }

  myql_instrument to_capi() const {
    myql_instrument o = myql::detail::blank_instrument(T_);
    o.kind = Tracker::capi_kind;
    o.barrier_dir = Dir == BarrierDirection::Up ? MYQL_BARRIER_UP : MYQL_BARRIER_DOWN;
    o.barrier_action = Action == BarrierAction::In ? MYQL_BARRIER_IN : MYQL_BARRIER_OUT;
    o.rebate_timing = Timing == RebateTiming::None    ? MYQL_REBATE_NONE
                      : Timing == RebateTiming::AtHit ? MYQL_REBATE_AT_HIT
                                                      : MYQL_REBATE_AT_MATURITY;
    o.barrier = barrier_;
    o.rebate = rebate_;
    o.rebate_rate = rate_;
    myql::detail::set_payoff(o, payoff_);
    myql::detail::attach_strikes(o, strikes_);
    return o;
  }

 private:
  StrikeContainer strikes_;
  double barrier_;
  double T_;
  double rate_;
  double rebate_;
  PayoffT payoff_;
};

template <RebateTiming Timing = RebateTiming::None>
using UpAndOutCall = BarrierOption<PayoffVanilla<OptionType::Call>, BarrierDirection::Up, BarrierAction::Out, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using UpAndOutPut = BarrierOption<PayoffVanilla<OptionType::Put>, BarrierDirection::Up, BarrierAction::Out, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using UpAndInCall = BarrierOption<PayoffVanilla<OptionType::Call>, BarrierDirection::Up, BarrierAction::In, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using UpAndInPut = BarrierOption<PayoffVanilla<OptionType::Put>, BarrierDirection::Up, BarrierAction::In, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using DownAndOutCall = BarrierOption<PayoffVanilla<OptionType::Call>, BarrierDirection::Down, BarrierAction::Out, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using DownAndOutPut = BarrierOption<PayoffVanilla<OptionType::Put>, BarrierDirection::Down, BarrierAction::Out, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using DownAndInCall = BarrierOption<PayoffVanilla<OptionType::Call>, BarrierDirection::Down, BarrierAction::In, Timing>;
template <RebateTiming Timing = RebateTiming::None>
using DownAndInPut = BarrierOption<PayoffVanilla<OptionType::Put>, BarrierDirection::Down, BarrierAction::In, Timing>;
