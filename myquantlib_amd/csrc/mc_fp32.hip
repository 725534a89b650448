// Instantiations of the path kernels with float path arithmetic.
#define MYQL_R float
#include "mc_inst.hpp"
