// Kernel specialisations for model XYNBSXSY, camera SCMOS, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBSXSY, SCMOS, PREC_FP32)
