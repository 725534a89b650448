// Jump policies (reference include/myql/models/asvj/policies/
// JumpPolicies.hpp:15-135) as tags carrying their parameter type.
#pragma once
#include <myql_capi.h>

#include <myql/models/asvj/data/ModelParams.hpp>

struct NoJumps {
  using Params = NoJumpParams;
  static constexpr int capi_id = MYQL_JUMP_NONE;
};
struct MertonJump {
  using Params = MertonParams;
  static constexpr int capi_id = MYQL_JUMP_MERTON;
};
struct KouJump {
  using Params = KouParams;
  static constexpr int capi_id = MYQL_JUMP_KOU;
};
