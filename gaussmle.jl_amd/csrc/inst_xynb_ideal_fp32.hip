// Kernel specialisations for model XYNB, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
// 11 px on 4 lanes is built in inst_sched_maxilp.hip.
GMLE_EXTERN_FIT(XYNB, IDEAL, PREC_FP32, 11, 4)
GMLE_DEFINE_FINDER(XYNB, IDEAL, PREC_FP32)
