// Kernel specialisations for model ASTIG, camera IDEAL, precision FP32.
#include "gmle_instances.hpp"
// 15 px on 4 lanes is built in inst_sched_ilp.hip.
GMLE_EXTERN_FIT(ASTIG, IDEAL, PREC_FP32, 15, 4)
GMLE_DEFINE_FINDER(ASTIG, IDEAL, PREC_FP32)
