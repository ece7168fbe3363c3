// Shared device/host helpers for the channel3D projection library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace c3d {

// Complex value aligned to its full size, so that LDS and global accesses
// of one element are single 16-byte (FP64) or 8-byte (FP32) operations.
template <typename real>
struct alignas(2 * sizeof(real)) cpx {
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

// XCD-aware logical workgroup coordinates.  MI355X dispatches workgroups
// round-robin over its 8 XCDs (hardware id b runs on XCD b % 8), each with
// its own L2.  The remap gives every XCD one contiguous range of logical
// ids, so tiles that are neighbours in logical order (and share halo rows)
// run on the same XCD.  A bijection on [0, n) for any grid size.
__device__ __forceinline__ void xcd_tile(unsigned& bx, unsigned& by, unsigned& bz) {
  constexpr unsigned kXcd = 8;
  const unsigned n = gridDim.x * gridDim.y * gridDim.z;
  const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned q = n / kXcd, r = n % kXcd, x = b % kXcd, slot = b / kXcd;
  const unsigned lid = x * q + (x < r ? x : r) + slot;
  bx = lid % gridDim.x;
  by = (lid / gridDim.x) % gridDim.y;
  bz = lid / (gridDim.x * gridDim.y);
}

// Whole-wavefront lane shift of an FP64 value through DPP (wave_shl:1, two
// 32-bit moves, no LDS): lane L receives the value of lane L+1.  Lane 63, and
// any lane whose source lane is inactive, receives 0, so the caller supplies
// those values itself.
__device__ __forceinline__ double wave_shl1(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x130, 0xf, 0xf, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Fortran column-major offset of 1-based (i,j,k) in an (n1,n2,*) array.
__host__ __device__ __forceinline__ size_t fidx(int i, int j, int k, int n1, int n2) {
  return (size_t)(i - 1) + (size_t)n1 * ((size_t)(j - 1) + (size_t)n2 * (size_t)(k - 1));
}

}  // namespace c3d
