// Payoff policies (reference include/myql/instruments/Payoffs.hpp:6-69).
// The payoff is evaluated on the GPU; these types carry the option type and
// the digital smoothing bandwidth (eps < 0: chosen automatically when
// Greeks are requested, as in the reference).
#pragma once
#include <myql_capi.h>

enum class OptionType { Call, Put };

template <OptionType OT>
struct PayoffVanilla {
  static constexpr OptionType Type = OT;
  static constexpr int capi_id = MYQL_PAYOFF_VANILLA;
  double smoothing_eps() const { return -1.0; }
};

template <OptionType OT>
struct PayoffCashOrNothing {
  static constexpr OptionType Type = OT;
  static constexpr int capi_id = MYQL_PAYOFF_CASH;
  double eps = -1.0;
  double smoothing_eps() const { return eps; }
};

template <OptionType OT>
struct PayoffAssetOrNothing {
  static constexpr OptionType Type = OT;
  static constexpr int capi_id = MYQL_PAYOFF_ASSET;
  double eps = -1.0;
  double smoothing_eps() const { return eps; }
};
