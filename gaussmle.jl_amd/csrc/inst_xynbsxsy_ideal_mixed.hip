// Kernel specialisations for model XYNBSXSY, camera IDEAL, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBSXSY, IDEAL, PREC_MIXED)
