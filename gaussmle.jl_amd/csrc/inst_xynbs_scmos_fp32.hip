// Kernel specialisations for model XYNBS, camera SCMOS, precision FP32.
#include "gmle_instances.hpp"
// 13 px on 4 lanes is built in inst_sched_ilp.hip.
GMLE_EXTERN_FIT(XYNBS, SCMOS, PREC_FP32, 13, 4)
GMLE_DEFINE_FINDER(XYNBS, SCMOS, PREC_FP32)
