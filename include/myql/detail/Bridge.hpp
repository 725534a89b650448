// Glue between the C++20 template layer and the C-ABI: one engine context
// per host thread, model / stepper marshalling, and the mapping of status
// codes onto the exceptions the reference throws (std::runtime_error for
// the SchemeNV condition, std::invalid_argument for bad arguments).
#pragma once
#include <myql_capi.h>

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

#include <myql/models/asvj/core/ASVJmodel.hpp>

namespace myql {

// Tag for the only generator the engine runs: xoshiro256++ with one stream
// per GPU lane (the reference's default RNG type).
struct xoshiro256pp {};

namespace detail {

// Context of the calling thread on HIP device $MYQL_DEVICE (default 0),
// created on first use and destroyed at thread exit.
class ThreadContext {
 public:
  ThreadContext() {
    const char *d = std::getenv("MYQL_DEVICE");
    ctx_ = myql_create(d ? std::atoi(d) : 0);
  }
  ~ThreadContext() {
    if (ctx_) myql_destroy(ctx_);
  }
  ThreadContext(const ThreadContext &) = delete;
  ThreadContext &operator=(const ThreadContext &) = delete;
  myql_ctx *get() {
    if (!ctx_) throw std::runtime_error("myql: no HIP device available for the Monte Carlo engine");
    return ctx_;
  }

 private:
  myql_ctx *ctx_ = nullptr;
};

inline myql_ctx *thread_context() {
  thread_local ThreadContext tc;
  return tc.get();
}

inline void check(int status, myql_ctx *ctx) {
  if (status == MYQL_OK) return;
  const std::string msg = myql_last_error(ctx);
  switch (status) {
    case MYQL_E_NV_FELLER:
      throw std::runtime_error(msg);
    case MYQL_E_INVALID_ARG:
      throw std::invalid_argument(msg);
    default:
      throw std::runtime_error("myql: " + msg);
  }
}

inline myql_heston_params to_capi(const HestonParams &h) {
  return myql_heston_params{h.kappa, h.theta, h.sigma, h.rho, h.v0};
}

template <typename JumpParams>
void set_jump(myql_model &m, const JumpParams &j) {
  if constexpr (std::is_same_v<JumpParams, MertonParams>) {
    m.jump[0] = j.lambda;
    m.jump[1] = j.mu;
    m.jump[2] = j.delta;
  } else if constexpr (std::is_same_v<JumpParams, KouParams>) {
    m.jump[0] = j.lambda;
    m.jump[1] = j.p_up;
    m.jump[2] = j.eta1;
    m.jump[3] = j.eta2;
  }
}

template <typename JumpParams>
inline constexpr int32_t jump_capi_id = std::is_same_v<JumpParams, MertonParams> ? MYQL_JUMP_MERTON
                                        : std::is_same_v<JumpParams, KouParams>  ? MYQL_JUMP_KOU
                                                                                 : MYQL_JUMP_NONE;

// Model alone -> C-ABI model (the Fourier pricer needs no stepper).
template <typename Model>
myql_model affine_model_to_capi(const Model &model) {
  myql_model m{};
  m.n_factors = Model::num_variance_factors;
  m.jump_kind = jump_capi_id<std::remove_cv_t<decltype(model.jump)>>;
  if constexpr (Model::num_variance_factors == 0) {
    m.vol = model.vol;
  } else if constexpr (Model::num_variance_factors == 1) {
    m.heston[0] = to_capi(model.heston);
  } else {
    m.heston[0] = to_capi(model.heston1);
    m.heston[1] = to_capi(model.heston2);
  }
  set_jump(m, model.jump);
  return m;
}

// Model + stepper -> C-ABI model and scheme pair, with the consistency
// checks the reference gets from its types.
template <typename Model, typename Stepper>
myql_model model_to_capi(const Model &model, int32_t schemes[2]) {
  static_assert(Model::num_variance_factors == Stepper::NumCirs,
                "the stepper's variance schemes do not match the model's factor count");
  static_assert(std::is_same_v<std::remove_cv_t<decltype(model.jump)>, typename Stepper::JumpP>,
                "the stepper's jump policy does not match the model's jump parameters");
  myql_model m{};
  m.n_factors = Model::num_variance_factors;
  m.jump_kind = Stepper::Jump::capi_id;
  if constexpr (Model::num_variance_factors == 0) {
    m.vol = model.vol;
  } else if constexpr (Model::num_variance_factors == 1) {
    m.heston[0] = to_capi(model.heston);
  } else {
    m.heston[0] = to_capi(model.heston1);
    m.heston[1] = to_capi(model.heston2);
  }
  set_jump(m, model.jump);
  schemes[0] = Stepper::Scheme1::capi_id;
  schemes[1] = Stepper::Scheme2::capi_id;
  return m;
}

}  // namespace detail
}  // namespace myql
