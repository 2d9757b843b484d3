// Kernel specialisations for model XYNBSXSY, camera SCMOS, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBSXSY, SCMOS, PREC_MIXED)
