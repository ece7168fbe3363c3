// Shared device/host helpers for the channel3D projection library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace c3d {

template <typename real>
struct cpx {
  real x, y;
};

template <typename real>
__device__ __forceinline__ cpx<real> cmk(real a, real b) {
  cpx<real> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename real>
__device__ __forceinline__ cpx<real> operator+(cpx<real> a, cpx<real> b) {
  return cmk(a.x + b.x, a.y + b.y);
}
template <typename real>
__device__ __forceinline__ cpx<real> operator-(cpx<real> a, cpx<real> b) {
  return cmk(a.x - b.x, a.y - b.y);
}
template <typename real>
__device__ __forceinline__ cpx<real> cmul(cpx<real> a, cpx<real> b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename real>
__device__ __forceinline__ cpx<real> cmulc(cpx<real> a, cpx<real> b) {
  return cmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Fortran column-major offset of 1-based (i,j,k) in an (n1,n2,*) array.
__host__ __device__ __forceinline__ size_t fidx(int i, int j, int k, int n1, int n2) {
  return (size_t)(i - 1) + (size_t)n1 * ((size_t)(j - 1) + (size_t)n2 * (size_t)(k - 1));
}

}  // namespace c3d
