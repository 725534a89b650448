// Caller of the runtime dispatcher (myql/dispatcher/*.hpp: build_model,
// build_instrument, build_european, price_mc, price_fourier, DispatchResult),
// linked to the GPU engine. Built twice: over the drop-in's own dispatcher
// (tests/cpp/dispatch_check_dropin, include/ alone on the include path) and,
// where the reference's headers are present, over the reference's
// dispatcher headers compiled where they lie with every other myql/ header
// from the drop-in (oracle/_ref/dispatch_check). price_mc / price_fourier
// take every (model, instrument) pair of AnyModel x AnyMCInstrument /
// AnyEuropeanInstrument, so the reference build instantiates the drop-in
// MonteCarloPricer and FourierPricer for all of them. Each case prints one
// JSON line; the Python binding of the same C-ABI must give the same
// numbers.
#include <myql/dispatcher/PricingDispatch.hpp>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

using namespace myql;
using namespace myql::dispatcher;

namespace {

void put_vec(const char *name, const std::vector<double> &v) {
  std::printf(", \"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]");
}

void emit(const std::string &name, const DispatchResult &r) {
  std::printf("{\"case\": \"%s\"", name.c_str());
  put_vec("prices", r.prices);
  put_vec("deltas", r.deltas);
  put_vec("gammas", r.gammas);
  for (size_t f = 0; f < r.vegas.size(); ++f) put_vec(f ? "vegas1" : "vegas0", r.vegas[f]);
  put_vec("thetas", r.thetas);
  put_vec("rhos", r.rhos);
  put_vec("prices_std_err", r.prices_std_err);
  put_vec("deltas_std_err", r.deltas_std_err);
  put_vec("gammas_std_err", r.gammas_std_err);
  for (size_t f = 0; f < r.vegas_std_err.size(); ++f)
    put_vec(f ? "vegas_std_err1" : "vegas_std_err0", r.vegas_std_err[f]);
  put_vec("thetas_std_err", r.thetas_std_err);
  put_vec("rhos_std_err", r.rhos_std_err);
  std::printf(", \"time_ms\": %.17g}\n", r.time_ms);
}

void emit_error(const std::string &name, const std::exception &e) {
  std::printf("{\"case\": \"%s\", \"error\": \"%s\"}\n", name.c_str(), e.what());
}

MonteCarloConfig mc(size_t paths, size_t steps) {
  MonteCarloConfig c;
  c.num_paths = paths;
  c.time_steps = steps;
  c.seed = 42;
  return c;
}

}  // namespace

int main() {
  const double S0 = 100.0, r = 0.02, q = 0.0;
  RuntimeModelParams mp;
  mp.heston1 = HestonParams{2.0, 0.04, 0.3, -0.7, 0.04};
  mp.heston2 = HestonParams{1.0, 0.03, 0.25, -0.5, 0.02};
  mp.merton = MertonParams{0.5, -0.1, 0.15};
  mp.kou = KouParams{1.0, 0.4, 10.0, 5.0};
  RuntimeModelParams bates = mp;
  bates.heston1 = HestonParams{2.0, 0.04, 1.0, -0.7, 0.04};

  RuntimeInstrumentParams strip;
  strip.strikes = {90.0, 100.0, 110.0};
  strip.maturity = 0.5;
  RuntimeInstrumentParams one;
  one.strikes = {100.0};
  one.maturity = 1.0;
  one.barrier_level = 130.0;

  try {
    emit("bates_nci_vanilla_call_full",
         price_mc<GreekMode::Full, SchemeNCI>(build_model("Bates", bates),
                                              build_instrument("Vanilla", "Call", strip), mc(1 << 14, 25),
                                              S0, r, q));
    emit("heston_nv_up_and_out_call",
         price_mc<GreekMode::None, SchemeNV>(build_model("Heston", mp), build_instrument("UpAndOut", "Call", one),
                                             mc(1 << 15, 64), S0, r, q));
    emit("double_heston_nv_arith_asian_call",
         price_mc<GreekMode::Essential, SchemeNV>(build_model("DoubleHeston", mp),
                                                  build_instrument("ArithAsian", "Call", one), mc(1 << 14, 32),
                                                  S0, r, q));
    emit("bates_kou_nv_lookback_floating_put",
         price_mc<GreekMode::Essential, SchemeNV>(build_model("BatesKou", mp),
                                                  build_instrument("LookbackFloating", "Put", one),
                                                  mc(1 << 14, 32), S0, r, q));
    emit("merton_cash_put_essential",
         price_mc<GreekMode::Essential, SchemeNV>(build_model("Merton", [&] {
                                                    RuntimeModelParams p = mp;
                                                    p.vol = 0.2;
                                                    return p;
                                                  }()),
                                                  build_instrument("CashOrNothing", "Put", strip), mc(1 << 14, 10),
                                                  S0, r, q));
    FourierEngine::Config fc;
    emit("fourier_heston_cash_call_full",
         price_fourier<GreekMode::Full>(build_model("Heston", mp), build_european("CashOrNothing", "Call", strip),
                                        fc, S0, r, q));
    emit("fourier_double_bates_kou_vanilla_put",
         price_fourier<GreekMode::Essential>(build_model("DoubleBatesKou", mp),
                                             build_european("Vanilla", "Put", strip), fc, S0, r, q));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "unexpected failure: %s\n", e.what());
    return 1;
  }
  // SchemeNV outside its domain (sigma^2 > 4 kappa theta) through the
  // dispatcher: the reference's exception text reaches the caller
  try {
    price_mc<GreekMode::None, SchemeNV>(build_model("Bates", bates), build_instrument("Vanilla", "Call", strip),
                                        mc(1 << 10, 10), S0, r, q);
    std::printf("{\"case\": \"scheme_nv_violation\"}\n");
  } catch (const std::runtime_error &e) {
    emit_error("scheme_nv_violation", e);
  }
  return 0;
}
