// Host buffers for one pricing call's per-strike outputs and their copy
// into the reference's result structs.
#pragma once
#include <myql_capi.h>

#include <array>
#include <cstddef>
#include <type_traits>
#include <vector>

#include <myql/core/PricingTypes.hpp>

namespace myql::detail {

class ResultArrays {
 public:
  enum { PRICE, PRICE_SE, DELTA, DELTA_SE, GAMMA, GAMMA_SE, VEGA0, VEGA1, VEGA_SE0, VEGA_SE1, THETA,
         THETA_SE, RHO, RHO_SE, VANNA0, VANNA1, VANNA_SE0, VANNA_SE1, CHARM, CHARM_SE, N_FIELDS };

  explicit ResultArrays(size_t n) {
    for (auto &f : f_) f.assign(n, 0.0);
  }

  myql_mc_result capi() {
    myql_mc_result r{};
    r.price = f_[PRICE].data();
    r.price_std_err = f_[PRICE_SE].data();
    r.delta = f_[DELTA].data();
    r.delta_std_err = f_[DELTA_SE].data();
    r.gamma = f_[GAMMA].data();
    r.gamma_std_err = f_[GAMMA_SE].data();
    r.vega[0] = f_[VEGA0].data();
    r.vega[1] = f_[VEGA1].data();
    r.vega_std_err[0] = f_[VEGA_SE0].data();
    r.vega_std_err[1] = f_[VEGA_SE1].data();
    r.theta = f_[THETA].data();
    r.theta_std_err = f_[THETA_SE].data();
    r.rho = f_[RHO].data();
    r.rho_std_err = f_[RHO_SE].data();
    r.vanna[0] = f_[VANNA0].data();
    r.vanna[1] = f_[VANNA1].data();
    r.vanna_std_err[0] = f_[VANNA_SE0].data();
    r.vanna_std_err[1] = f_[VANNA_SE1].data();
    r.charm = f_[CHARM].data();
    r.charm_std_err = f_[CHARM_SE].data();
    return r;
  }

  template <typename T>
  T get(int field) const {
    if constexpr (std::is_floating_point_v<T>) return f_[field][0];
    else return f_[field];
  }

  template <GreekMode Mode, typename T>
  void fill(FourierResult<Mode, T> &out) const {
    out.price = get<T>(PRICE);
    if constexpr (Mode != GreekMode::None) {
      out.delta = get<T>(DELTA);
      out.gamma = get<T>(GAMMA);
    }
    if constexpr (Mode == GreekMode::Full) {
      for (int i = 0; i < 2; ++i) {
        out.vega[i] = get<T>(VEGA0 + i);
        out.vanna[i] = get<T>(VANNA0 + i);
      }
      out.theta = get<T>(THETA);
      out.rho = get<T>(RHO);
      out.charm = get<T>(CHARM);
    }
  }

  template <GreekMode Mode, typename T>
  void fill(MonteCarloResult<Mode, T> &out) const {
    out.price = get<T>(PRICE);
    out.price_std_err = get<T>(PRICE_SE);
    if constexpr (Mode != GreekMode::None) {
      out.delta = get<T>(DELTA);
      out.delta_std_err = get<T>(DELTA_SE);
      out.gamma = get<T>(GAMMA);
      out.gamma_std_err = get<T>(GAMMA_SE);
    }
    if constexpr (Mode == GreekMode::Full) {
      for (int i = 0; i < 2; ++i) {
        out.vega[i] = get<T>(VEGA0 + i);
        out.vega_std_err[i] = get<T>(VEGA_SE0 + i);
        out.vanna[i] = get<T>(VANNA0 + i);
        out.vanna_std_err[i] = get<T>(VANNA_SE0 + i);
      }
      out.theta = get<T>(THETA);
      out.theta_std_err = get<T>(THETA_SE);
      out.rho = get<T>(RHO);
      out.rho_std_err = get<T>(RHO_SE);
      out.charm = get<T>(CHARM);
      out.charm_std_err = get<T>(CHARM_SE);
    }
  }

 private:
  std::array<std::vector<double>, N_FIELDS> f_;
};

template <GreekMode Mode>
inline constexpr int32_t greek_mode_id = Mode == GreekMode::None        ? MYQL_GREEK_NONE
                                         : Mode == GreekMode::Essential ? MYQL_GREEK_ESSENTIAL
                                                                        : MYQL_GREEK_FULL;

}  // namespace myql::detail
