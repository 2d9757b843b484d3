// Kernel specialisations for model ASTIG, camera SCMOS, precision FP32.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(ASTIG, SCMOS, PREC_FP32)
