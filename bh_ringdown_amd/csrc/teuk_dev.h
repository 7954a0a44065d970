// Device helpers shared by the libteuk kernel sources: global-memory views
// of device pointers, loads and stores by 32-bit byte offset, the wavefront
// LDS barrier, and the coefficient fields at one point.
#pragma once

#include "teuk_internal.h"

namespace teuk {
namespace {

// Every buffer a kernel touches is device global memory.  Pointers that
// arrive in kernel arguments are known to be global; pointers read from a
// table in memory (the batched kernels' descriptors) are not, and accesses
// through them would be flat instructions, which count against the scalar
// memory counter too, so that waiting for any scalar load also waits for
// every load in flight.  global_ptr states the address space.
#define TEUK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ TEUK_GLOBAL T *global_ptr(T *p) {
  return (TEUK_GLOBAL T *)p;
}

// Global accesses by a 32-bit byte offset from a uniform base: the address is
// an SGPR pair plus one VGPR (no 64-bit address arithmetic per access).  A
// solve's planes stay below 4 GiB (teuk_create checks N*M).
__device__ __forceinline__ double ld(const double *base, unsigned off) {
  return *global_ptr(reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + off));
}
__device__ __forceinline__ void st(double *base, unsigned off, double v) {
  *global_ptr(reinterpret_cast<double *>(reinterpret_cast<char *>(base) + off)) = v;
}
// A complex value (re, im) of a state: 16 bytes, one load or store (as a
// native two-double vector: HIP's double2 class has no address-space
// qualified copies).
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld2(const double *base, unsigned off) {
  const dvec2 v = *global_ptr(reinterpret_cast<const dvec2 *>(reinterpret_cast<const char *>(base) + off));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(double *base, unsigned off, double re, double im) {
  const dvec2 v = {re, im};
  *global_ptr(reinterpret_cast<dvec2 *>(reinterpret_cast<char *>(base) + off)) = v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

using teuk_pt::Coef;

// The nine coefficient fields at one point (kern.cu:132-158) from the
// column constants and the row values sin^2, cos and Y (teuk_point.h).
__device__ __forceinline__ Coef coef_point(const double (&K)[kNK], double sn2, double cs, double Y) {
  return teuk_pt::coef_point(K, sn2, cs, Y, [](double x) { return __dsqrt_rn(x); });
}

}  // namespace
}  // namespace teuk
