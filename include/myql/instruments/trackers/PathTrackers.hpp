// Path trackers (reference include/myql/instruments/trackers/
// PathTrackers.hpp:8-300) as tags; the tracking itself runs on the GPU.
#pragma once
#include <myql_capi.h>

struct TrackerEuropean {
  static constexpr int capi_kind = MYQL_INSTR_EUROPEAN;
};
struct TrackerGeoAsian {
  static constexpr int capi_kind = MYQL_INSTR_GEO_ASIAN;
};
struct TrackerArithAsian {
  static constexpr int capi_kind = MYQL_INSTR_ARITH_ASIAN;
};
struct TrackerLookback {
  static constexpr int capi_kind = MYQL_INSTR_LOOKBACK;
};
template <bool IsUpBarrier>
struct TrackerBarrier {
  static constexpr int capi_kind = MYQL_INSTR_BARRIER;
};
