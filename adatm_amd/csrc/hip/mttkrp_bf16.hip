/*
 * The memoized-MTTKRP kernels instantiated for FP32 values with BF16 factor rows
 * (templates in mttkrp_impl.hpp).
 */
#include "mttkrp_impl.hpp"

ADATM_MTTKRP_INSTANCE(bf16, float, bf16raw)
