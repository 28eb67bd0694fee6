// Host-only stand-in for <hip/hip_runtime.h>, used by tests/cpu/vjp_check.cpp
// to compile libdgx's pointwise device math (dgx_device.hpp, dgx_physics.hpp,
// dgx_vjp.hpp, dgx_dual.hpp) with g++.  Function qualifiers become no-ops and
// the hardware reciprocal estimate becomes an exact division (rcp() then
// refines it to the same value).  Test infrastructure only.
#pragma once
#include <cmath>
#include <cstddef>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline

typedef int hipError_t;
typedef void* hipStream_t;

inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
using std::fabs;
using std::sqrt;
