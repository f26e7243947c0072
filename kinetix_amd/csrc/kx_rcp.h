// FP64 reciprocal for the transport kernels: 1/y for finite nonzero y of moderate exponent
// (|y| and 1/|y| within the normal range), as v_rcp_f64 refined by two Newton steps, within
// 1 ulp of the quotient. Five dependent instructions instead of the IEEE division's ~10
// (scaling, Newton, fmas, fixup). No special cases: y = 0, infinities, NaN and results in the
// subnormal range are the caller's to avoid (the generated kernels use it for positive fit
// polynomials only). kx_selftest_rcp checks the 1-ulp bound on the device.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ double kx_rcp(double y)
{
  double r = __builtin_amdgcn_rcp(y);
  double e = fma(-y, r, 1.0);
  r = fma(r, e, r);
  e = fma(-y, r, 1.0);
  return fma(r, e, r);
}
