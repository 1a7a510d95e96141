/*
 * The memoized-MTTKRP kernels instantiated for FP32 values and factors
 * (templates in mttkrp_impl.hpp).
 */
#include "mttkrp_impl.hpp"

ADATM_MTTKRP_INSTANCE(fp32, float, float)
