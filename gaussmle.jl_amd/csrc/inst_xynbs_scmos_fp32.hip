// Kernel specialisations for model XYNBS, camera SCMOS, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBS, SCMOS, PREC_FP32)
