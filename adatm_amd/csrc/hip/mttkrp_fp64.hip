/*
 * The memoized-MTTKRP kernels instantiated for FP64 values and factors
 * (templates in mttkrp_impl.hpp).
 */
#include "mttkrp_impl.hpp"

ADATM_MTTKRP_INSTANCE(fp64, double, double)
