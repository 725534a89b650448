// Instantiations of the path kernels with double path arithmetic.
#define MYQL_R double
#include "mc_inst.hpp"
