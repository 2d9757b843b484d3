// Kernel specialisations for model XYNBS, camera SCMOS, precision MIXED.
#include "gmle_instances.hpp"
// The 13 px 4-lane iteration kernel is built in inst_sched_ilp.hip.
GMLE_EXTERN_KERNEL(XYNBS, SCMOS, PREC_MIXED, 13, 4)
GMLE_DEFINE_FINDER(XYNBS, SCMOS, PREC_MIXED)
