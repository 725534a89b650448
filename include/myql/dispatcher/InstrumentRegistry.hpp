// Drop-in runtime instrument registry (reference include/myql/dispatcher/
// InstrumentRegistry.hpp:14-260): the instrument aliases and variants the
// runtime dispatcher visits, the parameter bundle, and build_european /
// build_instrument by (type, option type) name. Same alias names, variant
// alternatives and order, parameter fields and names as the reference, so
// caller code is unchanged; the factories are tables of one constructor per
// (type, option type) pair.
#pragma once
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include <myql/instruments/Payoffs.hpp>
#include <myql/instruments/options/Asian.hpp>
#include <myql/instruments/options/Barrier.hpp>
#include <myql/instruments/options/European.hpp>
#include <myql/instruments/options/Lookback.hpp>

namespace myql::dispatcher {

// Europeans are strike strips; the exotics take one strike (or none).
using StandardStrip = std::vector<double>;

// ---- Europeans (Monte Carlo and Fourier) -----------------------------------
using VanillaCall = EuropeanOption<PayoffVanilla<OptionType::Call>, StandardStrip>;
using VanillaPut = EuropeanOption<PayoffVanilla<OptionType::Put>, StandardStrip>;
using CashCall = EuropeanOption<PayoffCashOrNothing<OptionType::Call>, StandardStrip>;
using CashPut = EuropeanOption<PayoffCashOrNothing<OptionType::Put>, StandardStrip>;
using AssetCall = EuropeanOption<PayoffAssetOrNothing<OptionType::Call>, StandardStrip>;
using AssetPut = EuropeanOption<PayoffAssetOrNothing<OptionType::Put>, StandardStrip>;

using AnyEuropeanInstrument = std::variant<VanillaCall, VanillaPut, CashCall, CashPut, AssetCall, AssetPut>;

// ---- Asians: fixed strike max(A - K, 0), floating max(S_T - A, 0) -----------
using GeoAsianFixedCall = AsianOption<TrackerGeoAsian, PayoffVanilla<OptionType::Call>, true>;
using GeoAsianFixedPut = AsianOption<TrackerGeoAsian, PayoffVanilla<OptionType::Put>, true>;
using GeoAsianFloatingCall = AsianOption<TrackerGeoAsian, PayoffVanilla<OptionType::Call>, false, double>;
using GeoAsianFloatingPut = AsianOption<TrackerGeoAsian, PayoffVanilla<OptionType::Put>, false, double>;
using ArithAsianFixedCall = AsianOption<TrackerArithAsian, PayoffVanilla<OptionType::Call>, true>;
using ArithAsianFixedPut = AsianOption<TrackerArithAsian, PayoffVanilla<OptionType::Put>, true>;
using ArithAsianFloatingCall = AsianOption<TrackerArithAsian, PayoffVanilla<OptionType::Call>, false, double>;
using ArithAsianFloatingPut = AsianOption<TrackerArithAsian, PayoffVanilla<OptionType::Put>, false, double>;

// ---- Lookbacks ----------------------------------------------------------------
using LookbackFixedCall = LookbackOption<PayoffVanilla<OptionType::Call>, true>;
using LookbackFixedPut = LookbackOption<PayoffVanilla<OptionType::Put>, true>;
using LookbackFloatingCall = LookbackOption<PayoffVanilla<OptionType::Call>, false, double>;
using LookbackFloatingPut = LookbackOption<PayoffVanilla<OptionType::Put>, false, double>;

// ---- Barriers (no rebate timing) ---------------------------------------------
using UpAndOutCallT = UpAndOutCall<RebateTiming::None>;
using UpAndOutPutT = UpAndOutPut<RebateTiming::None>;
using UpAndInCallT = UpAndInCall<RebateTiming::None>;
using UpAndInPutT = UpAndInPut<RebateTiming::None>;
using DownAndOutCallT = DownAndOutCall<RebateTiming::None>;
using DownAndOutPutT = DownAndOutPut<RebateTiming::None>;
using DownAndInCallT = DownAndInCall<RebateTiming::None>;
using DownAndInPutT = DownAndInPut<RebateTiming::None>;

using AnyMCInstrument =
    std::variant<VanillaCall, VanillaPut, CashCall, CashPut, AssetCall, AssetPut, GeoAsianFixedCall, GeoAsianFixedPut,
                 GeoAsianFloatingCall, GeoAsianFloatingPut, ArithAsianFixedCall, ArithAsianFixedPut,
                 ArithAsianFloatingCall, ArithAsianFloatingPut, LookbackFixedCall, LookbackFixedPut,
                 LookbackFloatingCall, LookbackFloatingPut, UpAndOutCallT, UpAndOutPutT, UpAndInCallT, UpAndInPutT,
                 DownAndOutCallT, DownAndOutPutT, DownAndInCallT, DownAndInPutT>;

struct RuntimeInstrumentParams {
  std::vector<double> strikes;  // Europeans: the strip; exotics: strikes[0] (fixed strike)
  double maturity;
  double barrier_level = 0.0;
  double rebate = 0.0;
  double rate = 0.0;  // discounting of an at-hit rebate
};

namespace detail {

template <typename V>
struct InstrumentEntry {
  const char *type, *option;
  V (*make)(const RuntimeInstrumentParams &);
};

inline double first_strike(const RuntimeInstrumentParams &p) { return p.strikes.empty() ? 0.0 : p.strikes[0]; }

// one row per (type, option type): a European from the strip
template <typename V, typename I>
V european(const RuntimeInstrumentParams &p) {
  return I(p.strikes, p.maturity);
}
// fixed-strike exotic from strikes[0]; floating-strike exotic with no strike
template <typename V, typename I>
V fixed(const RuntimeInstrumentParams &p) {
  return I(first_strike(p), p.maturity);
}
template <typename V, typename I>
V floating(const RuntimeInstrumentParams &p) {
  double unused = 0.0;
  return I(unused, p.maturity);
}
template <typename V, typename I>
V barrier(const RuntimeInstrumentParams &p) {
  return I(first_strike(p), p.barrier_level, p.maturity, p.rate, p.rebate);
}

template <typename V>
inline constexpr InstrumentEntry<V> kEuropeans[] = {
    {"Vanilla", "Call", european<V, VanillaCall>},        {"Vanilla", "Put", european<V, VanillaPut>},
    {"CashOrNothing", "Call", european<V, CashCall>},     {"CashOrNothing", "Put", european<V, CashPut>},
    {"AssetOrNothing", "Call", european<V, AssetCall>},   {"AssetOrNothing", "Put", european<V, AssetPut>},
};

using M = AnyMCInstrument;
inline constexpr InstrumentEntry<M> kExotics[] = {
    {"GeoAsian", "Call", fixed<M, GeoAsianFixedCall>},
    {"GeoAsian", "Put", fixed<M, GeoAsianFixedPut>},
    {"GeoAsianFloating", "Call", floating<M, GeoAsianFloatingCall>},
    {"GeoAsianFloating", "Put", floating<M, GeoAsianFloatingPut>},
    {"ArithAsian", "Call", fixed<M, ArithAsianFixedCall>},
    {"ArithAsian", "Put", fixed<M, ArithAsianFixedPut>},
    {"ArithAsianFloating", "Call", floating<M, ArithAsianFloatingCall>},
    {"ArithAsianFloating", "Put", floating<M, ArithAsianFloatingPut>},
    {"LookbackFixed", "Call", fixed<M, LookbackFixedCall>},
    {"LookbackFixed", "Put", fixed<M, LookbackFixedPut>},
    {"LookbackFloating", "Call", floating<M, LookbackFloatingCall>},
    {"LookbackFloating", "Put", floating<M, LookbackFloatingPut>},
    {"UpAndOut", "Call", barrier<M, UpAndOutCallT>},
    {"UpAndOut", "Put", barrier<M, UpAndOutPutT>},
    {"UpAndIn", "Call", barrier<M, UpAndInCallT>},
    {"UpAndIn", "Put", barrier<M, UpAndInPutT>},
    {"DownAndOut", "Call", barrier<M, DownAndOutCallT>},
    {"DownAndOut", "Put", barrier<M, DownAndOutPutT>},
    {"DownAndIn", "Call", barrier<M, DownAndInCallT>},
    {"DownAndIn", "Put", barrier<M, DownAndInPutT>},
};

template <typename V, size_t N>
const InstrumentEntry<V> *find(const InstrumentEntry<V> (&table)[N], const std::string &type,
                               const std::string &option) {
  for (const auto &e : table)
    if (type == e.type && option == e.option) return &e;
  return nullptr;
}

}  // namespace detail

// The European `type` ("Vanilla", "CashOrNothing", "AssetOrNothing") of
// `option_type` ("Call", "Put") on the strip params.strikes.
inline AnyEuropeanInstrument build_european(const std::string &type, const std::string &option_type,
                                            const RuntimeInstrumentParams &params) {
  if (const auto *e = detail::find(detail::kEuropeans<AnyEuropeanInstrument>, type, option_type))
    return e->make(params);
  throw std::invalid_argument("Unknown European instrument: " + type + " " + option_type);
}

// Any Monte Carlo instrument: the Europeans, then the Asians, lookbacks and
// barriers (strike params.strikes[0] where fixed).
inline AnyMCInstrument build_instrument(const std::string &type, const std::string &option_type,
                                        const RuntimeInstrumentParams &params) {
  if (const auto *e = detail::find(detail::kEuropeans<AnyMCInstrument>, type, option_type)) return e->make(params);
  if (const auto *e = detail::find(detail::kExotics, type, option_type)) return e->make(params);
  throw std::invalid_argument("Unknown instrument: " + type + " " + option_type);
}

}  // namespace myql::dispatcher
