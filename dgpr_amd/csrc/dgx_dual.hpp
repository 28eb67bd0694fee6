// dgx_dual.hpp -- forward-mode dual numbers: a value with D directional
// derivatives.  The pointwise physics (dgx_physics.hpp) is written for a
// generic scalar type, so running it on Dual<D> inputs seeded with unit
// directions gives D columns of its Jacobian at once; the reverse sweeps of
// the discrete adjoint (dgx_vjp.hpp, roe_flux_vjp) are checked against it
// (tests/cpu/vjp_check.cpp), and the adjoint's boundary-face kernel (a few
// face points per mesh) differentiates the boundary states and the boundary
// face physics with it.
#pragma once

#include <hip/hip_runtime.h>

#include "dgx_device.hpp"

namespace dgx
{

// the dual overloads below must not hide the double ones
using ::fabs;
using ::pow;
using ::sqrt;

template <int D>
struct Dual
{
  double v;
  double d[D];
  __device__ Dual() {}
  __device__ Dual(double x) : v(x)
  {
#pragma unroll
    for (int k = 0; k < D; ++k)
      d[k] = 0.0;
  }
};

template <int D>
__device__ __forceinline__ Dual<D> operator+(const Dual<D>& a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator-(const Dual<D>& a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator-(const Dual<D>& a)
{
  Dual<D> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = -a.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator*(const Dual<D>& a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator*(double a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = a * b.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator*(const Dual<D>& b, double a)
{
  return a * b;
}
template <int D>
__device__ __forceinline__ Dual<D> operator+(const Dual<D>& a, double b)
{
  Dual<D> r = a;
  r.v += b;
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator+(double b, const Dual<D>& a)
{
  return a + b;
}
template <int D>
__device__ __forceinline__ Dual<D> operator-(const Dual<D>& a, double b)
{
  Dual<D> r = a;
  r.v -= b;
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator-(double b, const Dual<D>& a)
{
  return (-a) + b;
}
template <int D>
__device__ __forceinline__ Dual<D>& operator+=(Dual<D>& a, const Dual<D>& b)
{
  a = a + b;
  return a;
}
template <int D>
__device__ __forceinline__ Dual<D>& operator-=(Dual<D>& a, const Dual<D>& b)
{
  a = a - b;
  return a;
}
template <int D>
__device__ __forceinline__ Dual<D>& operator*=(Dual<D>& a, double b)
{
  a = b * a;
  return a;
}
template <int D>
__device__ __forceinline__ bool operator<(const Dual<D>& a, const Dual<D>& b)
{
  return a.v < b.v;
}
// 1/x, with the same refined reciprocal as the residual kernels for the value
template <int D>
__device__ __forceinline__ Dual<D> rcp(const Dual<D>& x)
{
  Dual<D> r;
  r.v = rcp(x.v);
  const double s = -r.v * r.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = s * x.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> sqrt(const Dual<D>& x)
{
  Dual<D> r;
  r.v = ::sqrt(x.v);
  const double s = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = s * x.d[k];
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> fabs(const Dual<D>& x)
{
  return (x.v < 0.0) ? -x : x;
}
template <int D>
__device__ __forceinline__ bool operator>=(const Dual<D>& a, double b)
{
  return a.v >= b;
}
template <int D>
__device__ __forceinline__ Dual<D> operator/(const Dual<D>& a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator/(const Dual<D>& a, double b)
{
  Dual<D> r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = a.d[k] / b;
  return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator/(double a, const Dual<D>& b)
{
  Dual<D> r;
  r.v = a / b.v;
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = -r.v * b.d[k] / b.v;
  return r;
}
// x^e for x > 0
template <int D>
__device__ __forceinline__ Dual<D> pow(const Dual<D>& x, double e)
{
  Dual<D> r;
  r.v = ::pow(x.v, e);
  const double s = e * ::pow(x.v, e - 1.0);
#pragma unroll
  for (int k = 0; k < D; ++k)
    r.d[k] = s * x.d[k];
  return r;
}

}  // namespace dgx
