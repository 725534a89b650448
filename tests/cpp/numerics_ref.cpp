// Caller of the reference's own numerics (math/Numerics.hpp: Acklam's
// inverse normal normcdfinv_as241 and Giles' inverse Poisson CDF
// poissinvcdf; the header needs only <cmath>), built by oracle/Makefile's
// `ref` target from the reference's headers where they lie. It evaluates
// both on a fixed grid and prints one JSON object; tests/golden/
// gen_numerics_ref.py stores it as the fixture tests/golden/numerics_ref.json.
#include <myql/math/Numerics.hpp>

#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

// splitmix64, for reproducible pseudo-random grid points
uint64_t next(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit(uint64_t &s) { return static_cast<double>(next(s) >> 11) * 0x1.0p-53; }

void put_vec(const char *name, const std::vector<double> &v, bool first = false) {
  std::printf("%s\"%s\": [", first ? "" : ", ", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]");
}

}  // namespace

int main() {
  uint64_t s = 20261204;
  // inverse normal: both branches, their boundary (0.02425) and deep tails
  std::vector<double> p;
  for (int k = 1; k <= 300; k += 7) {
    p.push_back(std::pow(10.0, -k));
  }
  for (int k = 1; k <= 16; ++k) p.push_back(1.0 - std::pow(10.0, -k));
  for (double b : {0.02425, 0.97575}) {
    p.push_back(b);
    p.push_back(std::nextafter(b, 0.0));
    p.push_back(std::nextafter(b, 1.0));
  }
  for (int i = 0; i < 400; ++i) p.push_back(unit(s));
  std::vector<double> z;
  for (double x : p) z.push_back(Numerics::normcdfinv_as241(x));

  // inverse Poisson CDF: intensities across the summation / asymptotic /
  // tail branches, uniforms including the extremes
  std::vector<double> lam, u, n;
  const double lams[] = {1e-3, 0.05, 0.5, 1.0, 2.5, 3.999, 4.0, 4.001, 6.0, 9.5, 10.0, 12.0, 20.0, 31.7,
                         50.0, 100.0, 333.0, 1000.0, 1e4, 1e5};
  const double us[] = {1e-300, 1e-15, 1e-9, 1e-4, 0.01, 0.1, 0.5, 0.9, 0.99, 0.999999, 1.0 - 1e-12};
  for (double l : lams) {
    for (double x : us) {
      lam.push_back(l);
      u.push_back(x);
    }
    for (int i = 0; i < 60; ++i) {
      lam.push_back(l);
      u.push_back(unit(s));
    }
  }
  for (int i = 0; i < 1500; ++i) {
    lam.push_back(std::exp(std::log(1e-2) + unit(s) * (std::log(2e3) - std::log(1e-2))));
    u.push_back(unit(s));
  }
  for (size_t i = 0; i < lam.size(); ++i) n.push_back(Numerics::poissinvcdf(lam[i], u[i]));

  std::printf("{");
  put_vec("p", p, true);
  put_vec("z", z);
  put_vec("lam", lam);
  put_vec("u", u);
  put_vec("n", n);
  std::printf("}\n");
  return 0;
}
