// Drop-in stepper resolution (reference include/myql/dispatcher/
// StepperTraits.hpp:15-83): the stepper type a (model, instrument, variance
// scheme) triple prices with, for callers that name it. Same trait names
// and members as the reference. The runtime dispatcher itself does not
// need the type: the engine takes the model, schemes and tracker as run-time
// descriptors.
#pragma once
#include <type_traits>

#include <myql/models/asvj/core/ASVJmodel.hpp>
#include <myql/models/asvj/core/ASVJstepper.hpp>
#include <myql/models/asvj/data/ModelParams.hpp>
#include <myql/models/asvj/policies/JumpPolicies.hpp>
#include <myql/models/asvj/policies/VolSchemes.hpp>

namespace myql::dispatcher {

// jump parameter struct -> jump policy
template <typename ParamT>
struct JumpPolicyFor {
  using type = NoJumps;
};
template <>
struct JumpPolicyFor<MertonParams> {
  using type = MertonJump;
};
template <>
struct JumpPolicyFor<KouParams> {
  using type = KouJump;
};

// model -> variance factor count and jump parameter type
template <typename Model>
struct ModelTraits {
  static constexpr int num_cirs = Model::num_variance_factors;
  using jump_param_type = std::remove_cv_t<decltype(std::declval<Model>().jump)>;
};

template <typename Model, typename Instrument, typename DesiredVolScheme = SchemeNCI>
struct StepperFor {
  using Traits = ModelTraits<Model>;
  using JumpPolicy = typename JumpPolicyFor<typename Traits::jump_param_type>::type;
  using Tracker = typename Instrument::Tracker;
  static constexpr int NC = Traits::num_cirs;
  using V1 = std::conditional_t<(NC >= 1), DesiredVolScheme, NullVolScheme>;
  using V2 = std::conditional_t<(NC == 2), DesiredVolScheme, NullVolScheme>;
  using type = ASVJStepper<V1, V2, JumpPolicy, Tracker>;
};

}  // namespace myql::dispatcher
