// Kernel specialisations for model ASTIG, camera IDEAL, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(ASTIG, IDEAL, PREC_MIXED)
