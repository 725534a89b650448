// Stepper types (reference include/myql/models/asvj/core/ASVJstepper.hpp:
// 18-340). In the reference the stepper is the CPU hot loop; here it is a
// compile-time descriptor of the path the HIP kernels run: variance schemes
// of the (up to two) CIR factors, the jump law and the tracker.
#pragma once
#include <myql/instruments/trackers/PathTrackers.hpp>
#include <myql/models/asvj/core/ASVJmodel.hpp>
#include <myql/models/asvj/policies/JumpPolicies.hpp>
#include <myql/models/asvj/policies/VolSchemes.hpp>
#include <type_traits>

namespace myql::detail {
template <typename VS1, typename VS2>
inline constexpr int num_cirs =
    (std::is_same_v<VS1, NullVolScheme> && std::is_same_v<VS2, NullVolScheme>) ? 0
    : std::is_same_v<VS2, NullVolScheme>                                        ? 1
                                                                                : 2;
}  // namespace myql::detail

template <typename VolScheme1, typename VolScheme2 = NullVolScheme, typename JumpPolicy = NoJumps,
          typename TrackerPolicy = TrackerEuropean>
class ASVJStepper {
 public:
  static constexpr int NumCirs = myql::detail::num_cirs<VolScheme1, VolScheme2>;
  static_assert(NumCirs == 0 || !std::is_same_v<VolScheme1, NullVolScheme>,
                "VolScheme1 must be SchemeNV, SchemeExact or SchemeNCI");
  using Scheme1 = VolScheme1;
  using Scheme2 = VolScheme2;
  using Jump = JumpPolicy;
  using Policy = TrackerPolicy;
  using JumpP = typename JumpPolicy::Params;
  using ModelParams = std::conditional_t<NumCirs == 0, ZeroFactorModel<JumpP>,
                                         std::conditional_t<NumCirs == 1, SingleFactorModel<JumpP>,
                                                            DoubleFactorModel<JumpP>>>;
};

template <typename Tracker = TrackerEuropean>
using BlackScholesStepper = ASVJStepper<NullVolScheme, NullVolScheme, NoJumps, Tracker>;
template <typename Tracker = TrackerEuropean>
using MertonJDStepper = ASVJStepper<NullVolScheme, NullVolScheme, MertonJump, Tracker>;
template <typename Tracker = TrackerEuropean>
using KouJDStepper = ASVJStepper<NullVolScheme, NullVolScheme, KouJump, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using HestonStepper = ASVJStepper<VolScheme, NullVolScheme, NoJumps, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using BatesStepper = ASVJStepper<VolScheme, NullVolScheme, MertonJump, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using BatesKouStepper = ASVJStepper<VolScheme, NullVolScheme, KouJump, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using DoubleHestonStepper = ASVJStepper<VolScheme1, VolScheme2, NoJumps, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using DoubleBatesStepper = ASVJStepper<VolScheme1, VolScheme2, MertonJump, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using DoubleBatesKouStepper = ASVJStepper<VolScheme1, VolScheme2, KouJump, Tracker>;
