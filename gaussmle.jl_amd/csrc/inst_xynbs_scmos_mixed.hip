// Kernel specialisations for model XYNBS, camera SCMOS, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBS, SCMOS, PREC_MIXED)
