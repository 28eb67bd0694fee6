// dgx_dd.hpp -- double-double helpers of the basis conversion kernels
// (dgx_kernels.hip, dgx_general.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace dgx
{

// Double-double arithmetic (hi + lo) for the state basis conversions: the
// state itself is known to 1 ulp, and the residual's sensitivity to the state
// (the stiff acoustic part of the operator) would amplify a conversion error
// of a few ulps, so V = B3 U is formed with error-free transformations
// (Knuth two-sum, FMA two-product) and rounded once.
struct DD
{
  double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b)
{
  double s  = a + b;
  double bb = s - a;
  double e  = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ __forceinline__ DD dd_fast(double a, double b)
{
  double s = a + b;
  return {s, b - (s - a)};
}

// (x.hi + x.lo) * (c + clo) + acc, all in double-double
__device__ __forceinline__ DD dd_fma(DD x, double c, double clo, DD acc)
{
  double p = x.hi * c;
  double e = fma(x.hi, c, -p);
  e = fma(x.hi, clo, e);
  e = fma(x.lo, c, e);
  DD s = dd_two_sum(acc.hi, p);
  s.lo += acc.lo + e;
  return dd_fast(s.hi, s.lo);
}

}  // namespace dgx
