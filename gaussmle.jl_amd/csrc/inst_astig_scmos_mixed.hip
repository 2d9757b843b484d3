// Kernel specialisations for model ASTIG, camera SCMOS, precision MIXED.
#include "gmle_instances.hpp"
GMLE_DEFINE_FINDER(ASTIG, SCMOS, PREC_MIXED)
