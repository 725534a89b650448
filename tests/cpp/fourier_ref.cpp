// Caller of the reference's own Gil-Pelaez pricer (FourierPricer.hpp, its
// Boost-free closure), built by oracle/Makefile's `ref` target from the
// reference's headers where they lie (no drop-in headers on the include
// path). It prices a fixed set of European strips with GreekMode::Full and
// prints one JSON line per case; tests/golden/gen_fourier_ref.py turns the
// lines into the committed fixture tests/golden/fourier_ref.json.
//
// The reference's engine splits the integration range into
// max(8, 4 * omp_get_max_threads()) geometric chunks, so the numbers depend
// (at the adaptive-Simpson tolerance) on the thread count: the generator
// runs it with OMP_NUM_THREADS=16, which gives 64 chunks.
#include <myql/pricers/fourier/FourierPricer.hpp>

#include <cstdio>
#include <string>
#include <vector>

namespace {

const double S0 = 100.0, R = 0.03, Q = 0.01;
const std::vector<double> STRIKES = {60.0, 80.0, 95.0, 100.0, 105.0, 120.0, 150.0};

void put_vec(const char *name, const std::vector<double> &v) {
  std::printf(", \"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]");
}

template <typename Model, typename Payoff>
void run(const std::string &model_name, const Model &m, const char *payoff, const char *type, double T) {
  using Instr = EuropeanOption<Payoff, std::vector<double>>;
  FourierPricer<Model, Instr, GreekMode::Full> pricer(m, FourierEngine::Config{});
  const auto res = pricer.calculate(S0, R, Q, Instr(STRIKES, T));
  std::printf("{\"model\": \"%s\", \"payoff\": \"%s\", \"type\": \"%s\", \"T\": %.17g", model_name.c_str(), payoff,
              type, T);
  put_vec("strikes", STRIKES);
  put_vec("price", res.price);
  put_vec("delta", res.delta);
  put_vec("gamma", res.gamma);
  put_vec("vega0", res.vega[0]);
  put_vec("vega1", res.vega[1]);
  put_vec("theta", res.theta);
  put_vec("rho", res.rho);
  put_vec("vanna0", res.vanna[0]);
  put_vec("vanna1", res.vanna[1]);
  put_vec("charm", res.charm);
  std::printf("}\n");
}

template <typename Model>
void all_payoffs(const std::string &name, const Model &m, double T) {
  run<Model, PayoffVanilla<OptionType::Call>>(name, m, "vanilla", "call", T);
  run<Model, PayoffVanilla<OptionType::Put>>(name, m, "vanilla", "put", T);
  run<Model, PayoffCashOrNothing<OptionType::Call>>(name, m, "cash", "call", T);
  run<Model, PayoffCashOrNothing<OptionType::Put>>(name, m, "cash", "put", T);
  run<Model, PayoffAssetOrNothing<OptionType::Call>>(name, m, "asset", "call", T);
  run<Model, PayoffAssetOrNothing<OptionType::Put>>(name, m, "asset", "put", T);
}

}  // namespace

int main() {
  const HestonParams h1{2.0, 0.04, 0.3, -0.7, 0.04};
  const HestonParams h_hi{2.0, 0.04, 1.0, -0.7, 0.04};
  const HestonParams h2{1.0, 0.03, 0.25, -0.5, 0.02};
  const MertonParams mj{0.5, -0.1, 0.15};
  const KouParams kj{1.0, 0.4, 10.0, 5.0};
  for (double T : {0.25, 1.0, 2.0}) {
    all_payoffs("black_scholes", BlackScholesModel(0.25), T);
    all_payoffs("merton", MertonModel(0.2, mj), T);
    all_payoffs("kou", KouModel(0.2, kj), T);
    all_payoffs("heston", HestonModel(h1), T);
    all_payoffs("bates", BatesModel(h_hi, mj), T);
    all_payoffs("bates_kou", BatesKouModel(HestonParams{2.0, 0.04, 0.5, -0.7, 0.04}, kj), T);
    all_payoffs("double_heston", DoubleHestonModel(h1, h2), T);
    all_payoffs("double_bates", DoubleBatesModel(h1, h2, mj), T);
    all_payoffs("double_bates_kou", DoubleBatesKouModel(h1, h2, kj), T);
  }
  return 0;
}
