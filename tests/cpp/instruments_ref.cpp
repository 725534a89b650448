// Caller of the reference's own trackers and instruments
// (instruments/trackers/PathTrackers.hpp, instruments/options/{European,
// Asian,Lookback,Barrier}.hpp, instruments/Payoffs.hpp; no Boost), built by
// oracle/Makefile's `ref` target from the reference's headers where they
// lie. For every instrument below and a fixed set of log-price paths it runs
// the tracker in ASVJStepper's order (init at start_path, update after each
// step for path-dependent trackers, finalize over n * dt) and then
// calculate_to_buffer<GreekMode::Essential> (base, spot-up, spot-down
// payoffs), printing one JSON line per (instrument, path).
// tests/golden/gen_instruments_ref.py stores the lines as the fixture
// tests/golden/instruments_ref.json.
#include <myql/instruments/options/Asian.hpp>
#include <myql/instruments/options/Barrier.hpp>
#include <myql/instruments/options/European.hpp>
#include <myql/instruments/options/Lookback.hpp>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace {

const double S0 = 100.0, FD_BUMP = 1e-4, T = 1.0;
const int N_STEPS = 50;

uint64_t next(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit(uint64_t &s) { return static_cast<double>(next(s) >> 11) * 0x1.0p-53; }

// log-price paths: random walks of a few volatilities and drifts, and two
// that rise / fall steadily through the barriers
std::vector<std::vector<double>> make_paths() {
  std::vector<std::vector<double>> out;
  uint64_t s = 42;
  const double dt = T / N_STEPS;
  for (int p = 0; p < 6; ++p) {
    const double vol = 0.1 + 0.08 * p, drift = (p % 3 - 1) * 0.05;
    std::vector<double> x{std::log(S0)};
    for (int k = 0; k < N_STEPS; ++k) x.push_back(x.back() + drift * dt + vol * std::sqrt(dt) * (unit(s) - 0.5) * 3.4);
    out.push_back(x);
  }
  for (double slope : {0.3, -0.3}) {
    std::vector<double> x{std::log(S0)};
    for (int k = 0; k < N_STEPS; ++k) x.push_back(x.back() + slope / N_STEPS);
    out.push_back(x);
  }
  return out;
}

template <typename T>
struct StateFor : T::ExtraState {
  double logS = 0.0;
};

std::string num(double v) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}

void put_vec(const char *name, const std::vector<double> &v) {
  std::printf(", \"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]");
}

// `spec` is the instrument in the C-ABI's terms (JSON object body)
template <typename Instr>
void run(const std::string &name, const std::string &spec, const Instr &ins,
         const std::vector<std::vector<double>> &paths) {
  using Tr = typename Instr::Tracker;
  constexpr GreekMode Mode = GreekMode::Essential;
  const double dt = T / N_STEPS, h = S0 * FD_BUMP;
  for (size_t p = 0; p < paths.size(); ++p) {
    const auto &x = paths[p];
    StateFor<Tr> st;
    st.logS = x[0];
    const auto cfg = ins.template get_tracker_config<Mode>(S0, h, T);
    Tr::init(st, cfg, st.logS, dt);
    for (int k = 1; k <= N_STEPS; ++k) {
      st.logS = x[k];
      if constexpr (Tr::is_path_dependent) Tr::update(st, cfg, dt);
    }
    Tr::finalize(st, cfg, dt, N_STEPS * dt);
    std::vector<double> b, u, d;
    if constexpr (std::is_floating_point_v<typename Instr::ResultType>) {
      double bb = 0, uu = 0, dd = 0;
      ins.template calculate_to_buffer<Mode>(st, S0, h, bb, uu, dd);
      b = {bb};
      u = {uu};
      d = {dd};
    } else {
      typename Instr::ResultType bb(ins.size()), uu(ins.size()), dd(ins.size());
      ins.template calculate_to_buffer<Mode>(st, S0, h, bb, uu, dd);
      b = bb;
      u = uu;
      d = dd;
    }
    std::printf("{\"name\": \"%s\", \"path\": %zu, \"spec\": {%s}", name.c_str(), p, spec.c_str());
    put_vec("logS", x);
    put_vec("base", b);
    put_vec("up", u);
    put_vec("dn", d);
    std::printf("}\n");
  }
}

}  // namespace

int main() {
  const auto P = make_paths();
  const std::vector<double> K{80.0, 100.0, 120.0};
  using Call = PayoffVanilla<OptionType::Call>;
  using Put = PayoffVanilla<OptionType::Put>;
  PayoffCashOrNothing<OptionType::Call> cash;
  cash.eps = 2.0;
  PayoffAssetOrNothing<OptionType::Put> asset;
  asset.eps = 2.0;

  const std::string strip = "\"strikes\": [80, 100, 120], \"maturity\": 1";
  const std::string one = "\"strikes\": [100], \"maturity\": 1";
  run("european_call", "\"kind\": \"european\", \"payoff\": \"vanilla\", \"type\": \"call\", " + strip,
      EuropeanOption<Call, std::vector<double>>(K, T), P);
  run("european_put", "\"kind\": \"european\", \"payoff\": \"vanilla\", \"type\": \"put\", " + strip,
      EuropeanOption<Put, std::vector<double>>(K, T), P);
  run("european_cash_call",
      "\"kind\": \"european\", \"payoff\": \"cash\", \"type\": \"call\", \"payoff_eps\": 2, " + strip,
      EuropeanOption<PayoffCashOrNothing<OptionType::Call>, std::vector<double>>(K, T, cash), P);
  run("european_asset_put",
      "\"kind\": \"european\", \"payoff\": \"asset\", \"type\": \"put\", \"payoff_eps\": 2, " + strip,
      EuropeanOption<PayoffAssetOrNothing<OptionType::Put>, std::vector<double>>(K, T, asset), P);

  run("geo_fixed_call", "\"kind\": \"geo_asian\", \"type\": \"call\", \"fixed_strike\": true, " + one,
      AsianOption<TrackerGeoAsian, Call, true, double>(100.0, T), P);
  run("geo_floating_put", "\"kind\": \"geo_asian\", \"type\": \"put\", \"fixed_strike\": false, " + one,
      AsianOption<TrackerGeoAsian, Put, false, double>(100.0, T), P);
  run("geo_fixed_call_in_flight",
      "\"kind\": \"geo_asian\", \"type\": \"call\", \"fixed_strike\": true, \"past_avg_log\": " +
          num(std::log(95.0)) + ", \"t_elapsed\": 0.25, " + one,
      AsianOption<TrackerGeoAsian, Call, true, double>(100.0, T, std::log(95.0), 0.25), P);
  run("arith_fixed_put", "\"kind\": \"arith_asian\", \"type\": \"put\", \"fixed_strike\": true, " + one,
      AsianOption<TrackerArithAsian, Put, true, double>(100.0, T), P);
  run("arith_floating_call", "\"kind\": \"arith_asian\", \"type\": \"call\", \"fixed_strike\": false, " + one,
      AsianOption<TrackerArithAsian, Call, false, double>(100.0, T), P);
  run("arith_fixed_call_in_flight",
      "\"kind\": \"arith_asian\", \"type\": \"call\", \"fixed_strike\": true, \"past_avg_S\": 97, "
      "\"past_avg_log\": " + num(std::log(97.0)) + ", \"t_elapsed\": 0.5, " + one,
      AsianOption<TrackerArithAsian, Call, true, double>(100.0, T, 97.0, std::log(97.0), 0.5), P);

  run("lookback_fixed_call", "\"kind\": \"lookback\", \"type\": \"call\", \"fixed_strike\": true, " + one,
      LookbackOption<Call, true, double>(100.0, T), P);
  run("lookback_fixed_put", "\"kind\": \"lookback\", \"type\": \"put\", \"fixed_strike\": true, " + one,
      LookbackOption<Put, true, double>(100.0, T), P);
  run("lookback_floating_call", "\"kind\": \"lookback\", \"type\": \"call\", \"fixed_strike\": false, " + one,
      LookbackOption<Call, false, double>(0.0, T), P);
  run("lookback_floating_put", "\"kind\": \"lookback\", \"type\": \"put\", \"fixed_strike\": false, " + one,
      LookbackOption<Put, false, double>(0.0, T), P);
  run("lookback_floating_call_in_flight",
      "\"kind\": \"lookback\", \"type\": \"call\", \"fixed_strike\": false, \"hist_min_S\": 90, "
      "\"hist_max_S\": 115, " + one,
      LookbackOption<Call, false, double>(0.0, T, 90.0, 115.0), P);

  const std::string bar = "\"kind\": \"barrier\", \"strikes\": [100], \"maturity\": 1, ";
  run("up_out_call", bar + "\"type\": \"call\", \"dir\": \"up\", \"action\": \"out\", \"barrier\": 115",
      UpAndOutCall<>(100.0, 115.0, T), P);
  run("up_in_call", bar + "\"type\": \"call\", \"dir\": \"up\", \"action\": \"in\", \"barrier\": 115",
      UpAndInCall<>(100.0, 115.0, T), P);
  run("down_out_put", bar + "\"type\": \"put\", \"dir\": \"down\", \"action\": \"out\", \"barrier\": 90",
      DownAndOutPut<>(100.0, 90.0, T), P);
  run("down_in_put", bar + "\"type\": \"put\", \"dir\": \"down\", \"action\": \"in\", \"barrier\": 90",
      DownAndInPut<>(100.0, 90.0, T), P);
  run("up_out_call_rebate_at_hit",
      bar + "\"type\": \"call\", \"dir\": \"up\", \"action\": \"out\", \"barrier\": 115, \"rebate_timing\": "
            "\"at_hit\", \"rebate\": 2, \"rebate_rate\": 0.03",
      UpAndOutCall<RebateTiming::AtHit>(100.0, 115.0, T, 0.03, 2.0), P);
  run("down_in_put_rebate_at_maturity",
      bar + "\"type\": \"put\", \"dir\": \"down\", \"action\": \"in\", \"barrier\": 90, \"rebate_timing\": "
            "\"at_maturity\", \"rebate\": 1.5, \"rebate_rate\": 0.03",
      DownAndInPut<RebateTiming::AtMaturity>(100.0, 90.0, T, 0.03, 1.5), P);
  return 0;
}
