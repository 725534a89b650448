// Reference-style user code against the drop-in C++20 headers: the same
// includes, class templates and calculate() calls as code written for the
// reference library (compare benchmarks/test_smile_suite.cpp:60-80), linked
// with libmyql_hip.so. Prints one JSON object per case on stdout; the GPU
// tests compare them with the Python binding of the same C-ABI.
#include <cstdio>
#include <stdexcept>
#include <vector>

#include <myql/instruments/options/Asian.hpp>
#include <myql/instruments/options/Barrier.hpp>
#include <myql/instruments/options/European.hpp>
#include <myql/instruments/options/Lookback.hpp>
#include <myql/models/asvj/core/ASVJmodel.hpp>
#include <myql/models/asvj/core/ASVJstepper.hpp>
#include <myql/models/asvj/core/CoupledASVJStepper.hpp>
#include <myql/pricers/fourier/FourierPricer.hpp>
#include <myql/pricers/mlmc/MultiLevelMonteCarloPricer.hpp>
#include <myql/pricers/montecarlo/MonteCarloPricer.hpp>

static void print_vec(const char *name, const std::vector<double> &v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]%s", last ? "" : ", ");
}

int main() {
  const HestonParams heston{2.0, 0.04, 0.3, -0.7, 0.04};
  const double S0 = 100.0, r = 0.02, q = 0.0;

  {  // Heston, SchemeNV, European call, Essential Greeks
    HestonModel model(heston);
    MonteCarloConfig cfg;
    cfg.num_paths = 1 << 16;
    cfg.time_steps = 50;
    using Call = EuropeanOption<PayoffVanilla<OptionType::Call>>;
    MonteCarloPricer<HestonModel, HestonStepper<SchemeNV>, Call, GreekMode::Essential> pricer(model, cfg);
    const auto res = pricer.calculate(S0, r, q, Call(100.0, 1.0));
    std::printf("{\"case\": \"heston_nv_call\", \"price\": %.17g, \"price_std_err\": %.17g, \"delta\": %.17g, "
                "\"gamma\": %.17g, \"total_steps\": %zu}\n",
                res.price, res.price_std_err, res.delta, res.gamma, res.total_steps);
  }
  {  // Bates, SchemeNCI, strip of three calls, Full Greeks
    BatesModel model(HestonParams{2.0, 0.04, 1.0, -0.7, 0.04}, MertonParams{0.5, -0.1, 0.15});
    MonteCarloConfig cfg;
    cfg.num_paths = 1 << 14;
    cfg.time_steps = 25;
    using Strip = EuropeanOption<PayoffVanilla<OptionType::Call>, std::vector<double>>;
    MonteCarloPricer<BatesModel, BatesStepper<SchemeNCI>, Strip, GreekMode::Full> pricer(model, cfg);
    const auto res = pricer.calculate(S0, r, q, Strip({90.0, 100.0, 110.0}, 0.5));
    std::printf("{\"case\": \"bates_nci_strip\", ");
    print_vec("price", res.price);
    print_vec("delta", res.delta);
    print_vec("gamma", res.gamma);
    print_vec("vega0", res.vega[0]);
    print_vec("theta", res.theta);
    print_vec("rho", res.rho);
    print_vec("vanna0", res.vanna[0]);
    print_vec("charm", res.charm, true);
    std::printf("}\n");
  }
  {  // Heston NV up-and-out call, price only
    HestonModel model(heston);
    MonteCarloConfig cfg;
    cfg.num_paths = 1 << 15;
    cfg.time_steps = 64;
    using UOC = UpAndOutCall<>;
    MonteCarloPricer<HestonModel, HestonStepper<SchemeNV, UOC::Tracker>, UOC> pricer(model, cfg);
    const auto res = pricer.calculate(S0, r, q, UOC(100.0, 130.0, 1.0));
    std::printf("{\"case\": \"heston_uo_call\", \"price\": %.17g, \"price_std_err\": %.17g}\n", res.price,
                res.price_std_err);
  }
  {  // Heston NV arithmetic Asian fixed-strike call, Essential
    HestonModel model(heston);
    MonteCarloConfig cfg;
    cfg.num_paths = 1 << 15;
    cfg.time_steps = 32;
    using Asian = AsianOption<TrackerArithAsian, PayoffVanilla<OptionType::Call>, true>;
    MonteCarloPricer<HestonModel, HestonStepper<SchemeNV, TrackerArithAsian>, Asian, GreekMode::Essential> pricer(
        model, cfg);
    const auto res = pricer.calculate(S0, r, q, Asian(100.0, 1.0));
    std::printf("{\"case\": \"heston_asian_call\", \"price\": %.17g, \"delta\": %.17g}\n", res.price, res.delta);
  }
  {  // MLMC, Heston NV European call
    HestonModel model(heston);
    MLMCConfig cfg;
    cfg.epsilon = 0.05;
    cfg.L_min = 2;
    cfg.L_max = 4;
    cfg.N0 = 4096;
    cfg.base_steps = 2;
    using Call = EuropeanOption<PayoffVanilla<OptionType::Call>>;
    MultiLevelMonteCarloPricer<HestonModel, CoupledHestonStepper<SchemeNV>, Call> pricer(model, cfg);
    const auto res = pricer.calculate(S0, r, q, Call(100.0, 1.0));
    std::printf("{\"case\": \"heston_mlmc_call\", \"price\": %.17g, \"price_std_err\": %.17g, \"levels\": %zu}\n",
                res.price, res.price_std_err, res.paths_per_level.size());
  }
  {  // Gil-Pelaez: Bates strip with Full Greeks, and a cash-or-nothing put
    BatesModel model(HestonParams{2.0, 0.04, 1.0, -0.7, 0.04}, MertonParams{0.5, -0.1, 0.15});
    using Strip = EuropeanOption<PayoffVanilla<OptionType::Call>, std::vector<double>>;
    FourierPricer<BatesModel, Strip, GreekMode::Full> pricer(model);
    const auto res = pricer.calculate(S0, r, q, Strip({90.0, 100.0, 110.0}, 0.5));
    std::printf("{\"case\": \"bates_fourier_strip\", ");
    print_vec("price", res.price);
    print_vec("delta", res.delta);
    print_vec("vega0", res.vega[0]);
    print_vec("theta", res.theta, true);
    std::printf("}\n");
    using Digital = EuropeanOption<PayoffCashOrNothing<OptionType::Put>>;
    FourierPricer<BatesModel, Digital, GreekMode::Essential> dp(model);
    const auto d = dp.calculate(S0, r, q, Digital(95.0, 1.0));
    std::printf("{\"case\": \"bates_fourier_digital\", \"price\": %.17g, \"delta\": %.17g, \"gamma\": %.17g}\n",
                d.price, d.delta, d.gamma);
  }
  {  // SchemeNV outside its domain raises the reference's exception
    HestonModel model(HestonParams{2.0, 0.04, 1.0, -0.7, 0.04});
    MonteCarloConfig cfg;
    cfg.num_paths = 1024;
    using Call = EuropeanOption<PayoffVanilla<OptionType::Call>>;
    MonteCarloPricer<HestonModel, HestonStepper<SchemeNV>, Call> pricer(model, cfg);
    try {
      (void)pricer.calculate(S0, r, q, Call(100.0, 1.0));
      std::printf("{\"case\": \"nv_feller\", \"error\": null}\n");
    } catch (const std::runtime_error &e) {
      std::printf("{\"case\": \"nv_feller\", \"error\": \"%s\"}\n", e.what());
    }
  }
  return 0;
}
