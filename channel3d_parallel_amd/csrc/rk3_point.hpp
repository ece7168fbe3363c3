// Point functions of compute_rhs_u/v/w (src/equations.f90:24-423), shared
// by rk3.hip (FMA contraction off, divisions: bitwise the reference's
// operations) and rk3_fast.hip (contraction on, FAST = true: the spacing
// quotients as products with the tabulated reciprocals).  The translation
// unit that includes this header decides the contraction mode; the
// functions have internal linkage, so each unit keeps its own.
#pragma once

#include "kernels.hpp"

namespace c3d {
namespace {

// a / h as the reference writes it, or a * (1/h) with the tabulated reciprocal
template <bool FAST>
__device__ __forceinline__ double quot(double a, double h, double rh) {
  if constexpr (FAST) return a * rh;
  else return a / h;
}

// LDS tile geometry of k_rhs: 64 x 4 points per workgroup plus a one-cell
// halo in x and y, three z planes (k-1, k, k+1) kept as a ring.
constexpr int kTX = 64 + 2, kTY = 4 + 2, kTP = kTX * kTY;

// Field accessor of the point functions: U, V, W from the LDS tile.
struct TileAcc {
  const double* s;  // [field][slot][kTY][kTX]
  int i0, j0, k0;   // Fortran indices of tile element (1, 1) and of the ring's reference plane
  __device__ __forceinline__ int slot(int k) const { return ((k - k0) % 3 + 3) % 3; }
  __device__ __forceinline__ double at(int f, int i, int j, int k) const {
    return s[(f * 3 + slot(k)) * kTP + (j - j0 + 1) * kTX + (i - i0 + 1)];
  }
  __device__ __forceinline__ double U(int i, int j, int k) const { return at(0, i, j, k); }
  __device__ __forceinline__ double V(int i, int j, int k) const { return at(1, i, j, k); }
  __device__ __forceinline__ double W(int i, int j, int k) const { return at(2, i, j, k); }
};

// The x-direction spacing factors of the point functions (1/dxg, 1/dxf,
// nu/dxf, nu/dxg), read from the grid's arrays or, for a TileAccX, from the
// workgroup's LDS copy of the tile's x range i0-1 .. i0+64 (the same values:
// an LDS read in place of a global load per point).
enum { kXrdxg = 0, kXrdxf, kXnudxf, kXnudxg, kXT };
template <int T>
__device__ __forceinline__ const double* xtab_array(const DevGrid& g) {
  if constexpr (T == kXrdxg) return g.rdxg;
  else if constexpr (T == kXrdxf) return g.rdxf;
  else if constexpr (T == kXnudxf) return g.nudxf;
  else return g.nudxg;
}
struct TileAccX : TileAcc {
  const double* xt;  // [kXT][kTX]
};
template <int T>
__device__ __forceinline__ double xtab(const DevGrid& g, const TileAcc&, int i) {
  return xtab_array<T>(g)[i];
}
template <int T>
__device__ __forceinline__ double xtab(const DevGrid&, const TileAccX& A, int i) {
  return A.xt[T * kTX + (i - A.i0 + 1)];
}

// Plane k of U, V, W over the tile of A with its one-cell x/y halo, held in
// registers: element e = threadIdx.x + q * 256 (q = 0, 1) of the kTP tile
// cells; cells outside the arrays are zero.  load_plane issues the global
// loads, store_plane writes them to the plane's ring slot, so a kernel can
// overlap the loads of a later plane with the work on the current one.
struct PlaneRegs {
  double u[2], v[2], w[2];
};
static_assert(kTP <= 2 * 256, "two tile cells per thread of a 256-thread workgroup");

__device__ __forceinline__ void load_plane(const DevGrid& g, const double* __restrict__ U,
                                           const double* __restrict__ V,
                                           const double* __restrict__ W, const TileAcc& A, int k,
                                           PlaneRegs& r) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = threadIdx.x + q * 256;
    const int ii = A.i0 - 1 + e % kTX, jj = A.j0 - 1 + e / kTX;
    const bool ok = e < kTP && ii >= 1 && jj >= 1 && k >= 1;
    r.u[q] = (ok && ii <= nx && jj <= nyg && k <= nzg) ? U[fidx(ii, jj, k, nx, nyg)] : 0.0;
    r.v[q] = (ok && ii <= nxg && jj <= ny && k <= nzg) ? V[fidx(ii, jj, k, nxg, ny)] : 0.0;
    r.w[q] = (ok && ii <= nxg && jj <= nyg && k <= nz) ? W[fidx(ii, jj, k, nxg, nyg)] : 0.0;
  }
}

__device__ __forceinline__ void store_plane(double* tile, const TileAcc& A, int k,
                                            const PlaneRegs& r) {
  const int sl = A.slot(k);
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = threadIdx.x + q * 256;
    if (e < kTP) {
      tile[(0 * 3 + sl) * kTP + e] = r.u[q];
      tile[(1 * 3 + sl) * kTP + e] = r.v[q];
      tile[(2 * 3 + sl) * kTP + e] = r.w[q];
    }
  }
}

// Stages plane k of U, V, W (with the one-cell x/y halo around the tile of
// A) into its ring slot of the LDS tile; cells outside the arrays are zero.
__device__ __forceinline__ void stage_tile_plane(const DevGrid& g, const double* __restrict__ U,
                                                 const double* __restrict__ V,
                                                 const double* __restrict__ W, double* tile,
                                                 const TileAcc& A, int k, int nthreads) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  const int sl = A.slot(k);
  for (int e = threadIdx.x; e < kTP; e += nthreads) {
    const int ii = A.i0 - 1 + e % kTX, jj = A.j0 - 1 + e / kTX;
    const bool okx = ii >= 1, oky = jj >= 1;
    double u = 0.0, v = 0.0, w = 0.0;
    if (okx && oky && ii <= nx && jj <= nyg && k >= 1 && k <= nzg) u = U[fidx(ii, jj, k, nx, nyg)];
    if (okx && oky && ii <= nxg && jj <= ny && k >= 1 && k <= nzg) v = V[fidx(ii, jj, k, nxg, ny)];
    if (okx && oky && ii <= nxg && jj <= nyg && k >= 1 && k <= nz) w = W[fidx(ii, jj, k, nxg, nyg)];
    tile[(0 * 3 + sl) * kTP + e] = u;
    tile[(1 * 3 + sl) * kTP + e] = v;
    tile[(2 * 3 + sl) * kTP + e] = w;
  }
}

// compute_rhs_u, src/equations.f90:24-153; F bounds (2:nx-1, 2:nyg-1, 2:nzg-1)
template <bool FAST, typename Acc>
__device__ __forceinline__ double rhs_u_at(const DevGrid& g, const Acc& A,
                                         const double* __restrict__ H, double nu,
                                         double dPdx, int i, int j, int k) {
  const double uc1 = 0.5 * (A.U(i - 1, j, k) + A.U(i, j, k));
  const double uc2 = 0.5 * (A.U(i, j, k) + A.U(i + 1, j, k));
  const double a = quot<FAST>(-(uc2 * uc2 - uc1 * uc1), g.dxg[i], xtab<kXrdxg>(g, A, i));
  const double uv1 = (0.5 * (A.U(i, j, k) + A.U(i, j + 1, k))) * (0.5 * (A.V(i, j, k) + A.V(i + 1, j, k)));
  const double uv0 =
      (0.5 * (A.U(i, j - 1, k) + A.U(i, j, k))) * (0.5 * (A.V(i, j - 1, k) + A.V(i + 1, j - 1, k)));
  const double b = quot<FAST>(-(uv1 - uv0), g.dyf[j], g.rdyf[j]);
  const double uw1 = (0.5 * (A.U(i, j, k) + A.U(i, j, k + 1))) * (0.5 * (A.W(i, j, k) + A.W(i + 1, j, k)));
  const double uw0 =
      (0.5 * (A.U(i, j, k - 1) + A.U(i, j, k))) * (0.5 * (A.W(i, j, k - 1) + A.W(i + 1, j, k - 1)));
  const double c = quot<FAST>(-(uw1 - uw0), g.dzf[k], g.rdzf[k]);
  const double t1 = quot<FAST>(A.U(i, j + 1, k) - A.U(i, j, k), g.dyg[j], g.rdyg[j]);
  const double t1m = quot<FAST>(A.U(i, j, k) - A.U(i, j - 1, k), g.dyg[j - 1], g.rdyg[j - 1]);
  // 1.0/dx_3 and nu*1.0/dx_2 are index-only factors, tabulated (same
  // correctly rounded quotients as evaluating them here)
  const double visc = xtab<kXrdxg>(g, A, i) *
                          (xtab<kXnudxf>(g, A, i + 1) * (A.U(i + 1, j, k) - A.U(i, j, k)) -
                           xtab<kXnudxf>(g, A, i) * (A.U(i, j, k) - A.U(i - 1, j, k))) +
                      g.rdyf[j] * (nu * t1 - nu * t1m) +
                      g.rdzf[k] * (g.nudzg[k] * (A.U(i, j, k + 1) - A.U(i, j, k)) -
                                   g.nudzg[k - 1] * (A.U(i, j, k) - A.U(i, j, k - 1)));
  const double h = H ? H[fidx(i, j, k, g.nx, g.nyg)] : 1.0;
  double r = a;
  r = r + b;
  r = r + c;
  r = r + visc;
  r = r + dPdx * h;
  return r;
}

// compute_rhs_v, src/equations.f90:160-288; F bounds (2:nxg-1, 2:ny-1, 2:nzg-1)
template <bool FAST, typename Acc>
__device__ __forceinline__ double rhs_v_at(const DevGrid& g, const Acc& A,
                                         const double* __restrict__ H, double nu,
                                         double dPdy, int i, int j, int k) {
  const double vc1 = 0.5 * (A.V(i, j - 1, k) + A.V(i, j, k));
  const double vc2 = 0.5 * (A.V(i, j, k) + A.V(i, j + 1, k));
  const double a = quot<FAST>(-(vc2 * vc2 - vc1 * vc1), g.dyg[j], g.rdyg[j]);
  const double uv1 = (0.5 * (A.U(i, j, k) + A.U(i, j + 1, k))) * (0.5 * (A.V(i, j, k) + A.V(i + 1, j, k)));
  const double uv0 =
      (0.5 * (A.U(i - 1, j, k) + A.U(i - 1, j + 1, k))) * (0.5 * (A.V(i - 1, j, k) + A.V(i, j, k)));
  const double b = quot<FAST>(-(uv1 - uv0), g.dxf[i], xtab<kXrdxf>(g, A, i));
  const double vw1 = (0.5 * (A.V(i, j, k) + A.V(i, j, k + 1))) * (0.5 * (A.W(i, j, k) + A.W(i, j + 1, k)));
  const double vw0 =
      (0.5 * (A.V(i, j, k - 1) + A.V(i, j, k))) * (0.5 * (A.W(i, j, k - 1) + A.W(i, j + 1, k - 1)));
  const double c = quot<FAST>(-(vw1 - vw0), g.dzf[k], g.rdzf[k]);
  const double visc = xtab<kXrdxf>(g, A, i) *
                          (xtab<kXnudxg>(g, A, i) * (A.V(i + 1, j, k) - A.V(i, j, k)) -
                           xtab<kXnudxg>(g, A, i - 1) * (A.V(i, j, k) - A.V(i - 1, j, k))) +
                      g.rdyg[j] * (g.nudyf[j + 1] * (A.V(i, j + 1, k) - A.V(i, j, k)) -
                                   g.nudyf[j] * (A.V(i, j, k) - A.V(i, j - 1, k))) +
                      g.rdzf[k] * (g.nudzg[k] * (A.V(i, j, k + 1) - A.V(i, j, k)) -
                                   g.nudzg[k - 1] * (A.V(i, j, k) - A.V(i, j, k - 1)));
  const double h = H ? H[fidx(i, j, k, g.nxg, g.ny)] : 1.0;
  double r = a;
  r = r + b;
  r = r + c;
  r = r + visc;
  r = r + dPdy * h;
  return r;
}

// compute_rhs_w, src/equations.f90:295-423; F bounds (2:nxg-1, 2:nyg-1, 2:nz-1)
template <bool FAST, typename Acc>
__device__ __forceinline__ double rhs_w_at(const DevGrid& g, const Acc& A,
                                         const double* __restrict__ H, double nu,
                                         double dPdz, int i, int j, int k) {
  const double wc1 = 0.5 * (A.W(i, j, k - 1) + A.W(i, j, k));
  const double wc2 = 0.5 * (A.W(i, j, k) + A.W(i, j, k + 1));
  const double a = quot<FAST>(-(wc2 * wc2 - wc1 * wc1), g.dzg[k], g.rdzg[k]);
  const double uw1 = (0.5 * (A.U(i, j, k) + A.U(i, j, k + 1))) * (0.5 * (A.W(i, j, k) + A.W(i + 1, j, k)));
  const double uw0 =
      (0.5 * (A.U(i - 1, j, k) + A.U(i - 1, j, k + 1))) * (0.5 * (A.W(i - 1, j, k) + A.W(i, j, k)));
  const double b = quot<FAST>(-(uw1 - uw0), g.dxf[i], xtab<kXrdxf>(g, A, i));
  const double vw1 = (0.5 * (A.V(i, j, k) + A.V(i, j, k + 1))) * (0.5 * (A.W(i, j, k) + A.W(i, j + 1, k)));
  const double vw0 =
      (0.5 * (A.V(i, j - 1, k) + A.V(i, j - 1, k + 1))) * (0.5 * (A.W(i, j - 1, k) + A.W(i, j, k)));
  const double c = quot<FAST>(-(vw1 - vw0), g.dyf[j], g.rdyf[j]);
  const double t1 = quot<FAST>(A.W(i, j + 1, k) - A.W(i, j, k), g.dyg[j], g.rdyg[j]);
  const double t1m = quot<FAST>(A.W(i, j, k) - A.W(i, j - 1, k), g.dyg[j - 1], g.rdyg[j - 1]);
  const double visc = xtab<kXrdxf>(g, A, i) *
                          (xtab<kXnudxg>(g, A, i) * (A.W(i + 1, j, k) - A.W(i, j, k)) -
                           xtab<kXnudxg>(g, A, i - 1) * (A.W(i, j, k) - A.W(i - 1, j, k))) +
                      g.rdyf[j] * (nu * t1 - nu * t1m) +
                      g.rdzg[k] * (g.nudzf[k + 1] * (A.W(i, j, k + 1) - A.W(i, j, k)) -
                                   g.nudzf[k] * (A.W(i, j, k) - A.W(i, j, k - 1)));
  const double h = H ? H[fidx(i, j, k, g.nxg, g.nyg)] : 1.0;
  double r = a;
  r = r + b;
  r = r + c;
  r = r + visc;
  r = r + dPdz * h;
  return r;
}

}  // namespace
}  // namespace c3d
