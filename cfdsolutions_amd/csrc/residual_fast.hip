// Fast-arithmetic build of the residual-path kernels (CFD_ARITH_FAST): the
// same kernels as residual.hip, compiled with fused multiply-adds and with
// each shared denominator of the fluxes inverted once (flux.h,
// CFD_FAST_MATH).  FP64 results agree with the exact build to about 1e-15
// relative to the largest residual entry; the sums keep their order, so
// results are deterministic.  Launchers: namespace res_fast.
#pragma clang fp contract(fast)
#define CFD_FAST_MATH 1
#include "residual.hip"
