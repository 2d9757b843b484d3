// Kernel specialisations for model XYNB, camera IDEAL, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNB, IDEAL, PREC_MIXED)
