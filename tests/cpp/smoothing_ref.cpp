// Caller of the reference's Boost-free variance helpers
// (models/asvj/data/ModelParams.hpp:54-80) on the reference's own model
// containers (models/asvj/core/ASVJmodel.hpp), built by oracle/Makefile's
// `ref` target from the reference's headers where they lie. Per model and
// maturity it prints the expected average variance and the automatic
// digital-smoothing bandwidth of MonteCarloPricer.hpp:31-74:
//   var = diffusion part (vol^2, or the Heston helper per factor) + jump part
//   eps = sqrt(var) * scale * sqrt(T) / 10, scale = K (scalar strike) or S0
//         (strike strip).
// The helpers are the reference's; that composition is in
// MonteCarloPricer.hpp, whose closure needs Boost, so it is restated below
// (the `requires` dispatch of MonteCarloPricer.hpp:33-47).
// tests/golden/gen_smoothing_ref.py writes tests/golden/smoothing_ref.json.
#include <myql/models/asvj/core/ASVJmodel.hpp>

#include <cmath>
#include <cstdio>
#include <initializer_list>

namespace {

const double S0 = 100.0, K_SCALAR = 110.0;

template <typename Model>
double average_variance(const Model &m, double T) {
  using namespace myql::models::variance;
  double v = 0.0;
  if constexpr (requires { m.vol; }) {
    v += expected_average_variance(m.vol, T);
  } else if constexpr (requires { m.heston; }) {
    v += expected_average_variance(m.heston, T);
  } else if constexpr (requires { m.heston1; }) {
    v += expected_average_variance(m.heston1, T);
    v += expected_average_variance(m.heston2, T);
  }
  return v + jump_variance(m.jump);
}

template <typename Model>
void row(const char *name, const Model &m, double T) {
  const double var = average_variance(m, T);
  const double sig = std::sqrt(var);
  std::printf("{\"model\": \"%s\", \"T\": %.17g, \"avg_var\": %.17g, \"eps_scalar\": %.17g, \"eps_strip\": %.17g}\n",
              name, T, var, sig * K_SCALAR * std::sqrt(T) / 10.0, sig * S0 * std::sqrt(T) / 10.0);
}

}  // namespace

int main() {
  const HestonParams h1{2.0, 0.04, 0.3, -0.7, 0.04};
  const HestonParams h_hi{2.0, 0.04, 1.0, -0.7, 0.04};
  const HestonParams h2{1.0, 0.03, 0.25, -0.5, 0.02};
  const HestonParams h_bk{2.0, 0.04, 0.5, -0.7, 0.04};
  const HestonParams h_flat{1e-9, 0.05, 0.2, -0.3, 0.03};  // kappa T <= 1e-8: the v0 branch
  const MertonParams mj{0.5, -0.1, 0.15};
  const KouParams kj{1.0, 0.4, 10.0, 5.0};
  for (double T : {0.1, 0.5, 1.0, 2.0}) {
    row("black_scholes", BlackScholesModel(0.25), T);
    row("merton", MertonModel(0.2, mj), T);
    row("kou", KouModel(0.2, kj), T);
    row("heston", HestonModel(h1), T);
    row("bates", BatesModel(h_hi, mj), T);
    row("bates_kou", BatesKouModel(h_bk, kj), T);
    row("double_heston", DoubleHestonModel(h1, h2), T);
    row("double_bates", DoubleBatesModel(h1, h2, mj), T);
    row("double_bates_kou", DoubleBatesKouModel(h1, h2, kj), T);
    row("heston_flat", HestonModel(h_flat), T);
  }
  return 0;
}
