// Coupled fine/coarse stepper types for MLMC (reference include/myql/
// models/asvj/core/CoupledASVJStepper.hpp:20-615), as descriptors of the
// coupled path the HIP level kernels run.
#pragma once
#include <myql/models/asvj/core/ASVJstepper.hpp>

template <typename VolScheme1, typename VolScheme2 = NullVolScheme, typename JumpPolicy = NoJumps,
          typename TrackerPolicy = TrackerEuropean>
class CoupledASVJStepper : public ASVJStepper<VolScheme1, VolScheme2, JumpPolicy, TrackerPolicy> {};

template <typename Tracker = TrackerEuropean>
using CoupledBlackScholesStepper = CoupledASVJStepper<NullVolScheme, NullVolScheme, NoJumps, Tracker>;
template <typename Tracker = TrackerEuropean>
using CoupledMertonJDStepper = CoupledASVJStepper<NullVolScheme, NullVolScheme, MertonJump, Tracker>;
template <typename Tracker = TrackerEuropean>
using CoupledKouJDStepper = CoupledASVJStepper<NullVolScheme, NullVolScheme, KouJump, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using CoupledHestonStepper = CoupledASVJStepper<VolScheme, NullVolScheme, NoJumps, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using CoupledBatesStepper = CoupledASVJStepper<VolScheme, NullVolScheme, MertonJump, Tracker>;
template <typename VolScheme, typename Tracker = TrackerEuropean>
using CoupledBatesKouStepper = CoupledASVJStepper<VolScheme, NullVolScheme, KouJump, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using CoupledDoubleHestonStepper = CoupledASVJStepper<VolScheme1, VolScheme2, NoJumps, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using CoupledDoubleBatesStepper = CoupledASVJStepper<VolScheme1, VolScheme2, MertonJump, Tracker>;
template <typename VolScheme1, typename VolScheme2, typename Tracker = TrackerEuropean>
using CoupledDoubleBatesKouStepper = CoupledASVJStepper<VolScheme1, VolScheme2, KouJump, Tracker>;
