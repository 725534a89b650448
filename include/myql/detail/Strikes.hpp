// Strike containers of the C++20 layer: double (one option) or
// std::vector<double> (a strip priced on the same paths).
#pragma once
#include <myql_capi.h>

#include <type_traits>
#include <vector>

namespace myql::detail {

template <typename StrikeContainer>
inline constexpr bool is_scalar_strike = std::is_floating_point_v<StrikeContainer>;

// Point the C-ABI instrument at the strikes held by the C++ instrument.
template <typename StrikeContainer>
void attach_strikes(myql_instrument &out, const StrikeContainer &K) {
  if constexpr (is_scalar_strike<StrikeContainer>) {
    out.strikes = &K;
    out.n_strikes = 1;
    out.is_strip = 0;
  } else {
    out.strikes = K.data();
    out.n_strikes = static_cast<int64_t>(K.size());
    out.is_strip = 1;
  }
}

template <typename PayoffT>
void set_payoff(myql_instrument &out, const PayoffT &p) {
  out.payoff = PayoffT::capi_id;
  out.option_type = PayoffT::Type == decltype(PayoffT::Type)::Call ? MYQL_CALL : MYQL_PUT;
  out.payoff_eps = p.smoothing_eps();
}

inline myql_instrument blank_instrument(double T) {
  myql_instrument o{};
  o.maturity = T;
  o.payoff_eps = -1.0;
  return o;
}

}  // namespace myql::detail
