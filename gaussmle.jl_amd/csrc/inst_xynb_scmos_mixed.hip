// Kernel specialisations for model XYNB, camera SCMOS, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(XYNB, SCMOS, PREC_MIXED)
