// Kernel specialisations for model XYNBS, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
// 21 px on 8 lanes is built in inst_sched_maxilp.hip.
GMLE_EXTERN_FIT(XYNBS, IDEAL, PREC_FP32, 21, 8)
GMLE_DEFINE_FINDER(XYNBS, IDEAL, PREC_FP32)
