// Kernel specialisations for model XYNBS, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBS, IDEAL, PREC_FP32)
