// Kernel specialisations for model XYNBSXSY, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBSXSY, IDEAL, PREC_FP32)
