// Kernel specialisations for model XYNB, camera SCMOS, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNB, SCMOS, PREC_FP32)
