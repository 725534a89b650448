// Model containers (reference include/myql/models/asvj/core/ASVJmodel.hpp:
// 14-87): zero, one or two CIR variance factors plus a jump law.
#pragma once
#include <myql/models/asvj/data/ModelParams.hpp>

template <typename JumpParamType = NoJumpParams>
struct ZeroFactorModel {
  static constexpr int num_variance_factors = 0;
  double vol;
  JumpParamType jump;
  ZeroFactorModel(double v, const JumpParamType &j = JumpParamType{}) : vol(v), jump(j) {}
};

template <typename JumpParamType>
struct SingleFactorModel {
  static constexpr int num_variance_factors = 1;
  HestonParams heston;
  JumpParamType jump;
  SingleFactorModel(const HestonParams &h, const JumpParamType &j = JumpParamType())
      : heston(h), jump(j) {}
};

// two independent CIR factors sharing one stock Brownian motion
template <typename JumpParamType>
struct DoubleFactorModel {
  static constexpr int num_variance_factors = 2;
  HestonParams heston1;
  HestonParams heston2;
  JumpParamType jump;
  DoubleFactorModel(const HestonParams &h1, const HestonParams &h2,
                    const JumpParamType &j = JumpParamType())
      : heston1(h1), heston2(h2), jump(j) {}
};

using BlackScholesModel = ZeroFactorModel<NoJumpParams>;
using MertonModel = ZeroFactorModel<MertonParams>;
using KouModel = ZeroFactorModel<KouParams>;
using HestonModel = SingleFactorModel<NoJumpParams>;
using BatesModel = SingleFactorModel<MertonParams>;
using BatesKouModel = SingleFactorModel<KouParams>;
using DoubleHestonModel = DoubleFactorModel<NoJumpParams>;
using DoubleBatesModel = DoubleFactorModel<MertonParams>;
using DoubleBatesKouModel = DoubleFactorModel<KouParams>;
