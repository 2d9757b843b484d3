// Kernel specialisations for model ASTIG, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(ASTIG, IDEAL, PREC_FP32)
