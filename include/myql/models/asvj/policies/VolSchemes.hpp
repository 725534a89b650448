// Variance discretisation schemes (reference include/myql/models/asvj/
// policies/VolSchemes.hpp:19-377). Here they are tags: the schemes run in
// the HIP path kernels and a stepper type only selects them.
#pragma once
#include <myql_capi.h>

struct NullVolScheme {
  static constexpr int capi_id = MYQL_SCHEME_NULL;
};
// Ninomiya-Victoir; requires sigma^2 <= 4 kappa theta (std::runtime_error)
struct SchemeNV {
  static constexpr int capi_id = MYQL_SCHEME_NV;
};
// exact Poisson-gamma mixture (variable draw count)
struct SchemeExact {
  static constexpr int capi_id = MYQL_SCHEME_EXACT;
};
// Poisson mixture with tabulated chi-squared inverse
struct SchemeNCI {
  static constexpr int capi_id = MYQL_SCHEME_NCI;
};
