// Kernel specialisations for model XYNBS, camera IDEAL, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNBS, IDEAL, PREC_MIXED)
