// Drop-in runtime model registry (reference include/myql/dispatcher/
// ModelRegistry.hpp:13-74): the variant of the nine ASVJ models, the
// parameter bundle a caller fills at run time, and build_model by name.
// Same variant alternatives (in the same order), field names and model
// names as the reference, so caller code that switches on them, or visits
// the variant, is unchanged.
#pragma once
#include <stdexcept>
#include <string>
#include <variant>

#include <myql/models/asvj/core/ASVJmodel.hpp>

namespace myql::dispatcher {

using AnyModel = std::variant<BlackScholesModel, MertonModel, KouModel, HestonModel, BatesModel, BatesKouModel,
                              DoubleHestonModel, DoubleBatesModel, DoubleBatesKouModel>;

// Parameters of any of the nine models; each model reads the fields its
// factors and jump law need.
struct RuntimeModelParams {
  HestonParams heston1{};
  HestonParams heston2{};
  double vol = 0.0;  // zero-factor models
  MertonParams merton{};
  KouParams kou{};
};

namespace detail {
using ModelMaker = AnyModel (*)(const RuntimeModelParams &);
struct ModelEntry {
  const char *name;
  ModelMaker make;
};
// name -> constructor, one row per model
inline constexpr ModelEntry kModels[] = {
    {"BlackScholes", [](const RuntimeModelParams &p) -> AnyModel { return BlackScholesModel(p.vol); }},
    {"Merton", [](const RuntimeModelParams &p) -> AnyModel { return MertonModel(p.vol, p.merton); }},
    {"Kou", [](const RuntimeModelParams &p) -> AnyModel { return KouModel(p.vol, p.kou); }},
    {"Heston", [](const RuntimeModelParams &p) -> AnyModel { return HestonModel(p.heston1); }},
    {"Bates", [](const RuntimeModelParams &p) -> AnyModel { return BatesModel(p.heston1, p.merton); }},
    {"BatesKou", [](const RuntimeModelParams &p) -> AnyModel { return BatesKouModel(p.heston1, p.kou); }},
    {"DoubleHeston",
     [](const RuntimeModelParams &p) -> AnyModel { return DoubleHestonModel(p.heston1, p.heston2); }},
    {"DoubleBates",
     [](const RuntimeModelParams &p) -> AnyModel { return DoubleBatesModel(p.heston1, p.heston2, p.merton); }},
    {"DoubleBatesKou",
     [](const RuntimeModelParams &p) -> AnyModel { return DoubleBatesKouModel(p.heston1, p.heston2, p.kou); }},
};
}  // namespace detail

// The model named `model_type`; std::invalid_argument for an unknown name.
inline AnyModel build_model(const std::string &model_type, const RuntimeModelParams &params) {
  for (const auto &e : detail::kModels)
    if (model_type == e.name) return e.make(params);
  throw std::invalid_argument("Unknown model type: " + model_type);
}

}  // namespace myql::dispatcher
