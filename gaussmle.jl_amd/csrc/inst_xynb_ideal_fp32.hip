// Kernel specialisations for model XYNB, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNB, IDEAL, PREC_FP32)
