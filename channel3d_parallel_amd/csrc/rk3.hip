// RK3 advection-diffusion kernels (src/equations.f90, src/time_integration.f90,
// src/mass_flow.f90) for gfx950, FP64.
//
// Each right-hand side is evaluated pointwise: the reference's whole-array
// passes through term_1/term_2 reduce, at every interior point, to the
// expressions below (same operands, same operation order, same order of
// accumulation: convective x, y, z, viscous, pressure gradient).  FMA
// contraction is disabled so the results round exactly as written.
#include <cstdlib>

#include "kernels.hpp"

#pragma clang fp contract(off)
#include "rk3_point.hpp"

namespace c3d {

namespace {

constexpr int kB = 256;

#define U_(i, j, k) U[fidx(i, j, k, g.nx, g.nyg)]
#define V_(i, j, k) V[fidx(i, j, k, g.nxg, g.ny)]
#define W_(i, j, k) W[fidx(i, j, k, g.nxg, g.nyg)]

// compute_rhs_u/v/w evaluated together.  Ranges: F_u (2:nx-1, 2:nyg-1,
// 2:nzg-1), F_v (2:nxg-1, 2:ny-1, 2:nzg-1), F_w (2:nxg-1, 2:nyg-1, 2:nz-1);
// the grid covers their union.  Each workgroup owns a 64 x 4 column of
// points and marches KC z planes; U, V, W around it are staged in LDS as a
// ring of three planes with a one-cell x/y halo, so every neighbour read of
// the point functions is an LDS read and each field value is fetched from
// HBM about once.  Out-of-array halo cells are zero and only feed outputs
// outside the ranges above.
template <int KC>
__global__ __launch_bounds__(kB) void k_rhs(DevGrid g, const double* __restrict__ U,
                                            const double* __restrict__ V,
                                            const double* __restrict__ W,
                                            const double* __restrict__ Hu,
                                            const double* __restrict__ Hv,
                                            const double* __restrict__ Hw, double nu, double dPdx,
                                            double dPdy, double dPdz, double* __restrict__ Fu,
                                            double* __restrict__ Fv, double* __restrict__ Fw) {
  __shared__ double tile[3 * 3 * kTP];
  __shared__ double xt[kXT * kTX];
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);  // neighbouring tiles on one XCD (shared halo rows)
  const int i0 = bx * 64 + 2, j0 = by * 4 + 2, kb = bz * KC + 2;
  // j is the same on all lanes of a wavefront (one row of 64 x positions)
  const int i = i0 + (threadIdx.x & 63);
  const int j = __builtin_amdgcn_readfirstlane(j0 + (int)(threadIdx.x >> 6));
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  // LDS copy of the tile's x spacing factors (the same table values, so the
  // result stays bitwise); the first plane's barrier makes it visible
  for (int e = threadIdx.x; e < kXT * kTX; e += kB) {
    const int t = e / kTX, ii = min(i0 - 1 + e % kTX, nx + 1);
    const double* src = t == kXrdxg ? g.rdxg : t == kXrdxf ? g.rdxf : t == kXnudxf ? g.nudxf : g.nudxg;
    xt[e] = src[ii];
  }
  TileAccX A{{tile, i0, j0, kb}, xt};
  auto stage = [&](int k) { stage_tile_plane(g, U, V, W, tile, A, k, kB); };
  stage(kb - 1);
  stage(kb);
  for (int kk = 0; kk < KC; kk++) {
    const int k = kb + kk;
    if (k > nzg - 1) break;  // uniform across the workgroup
    stage(k + 1);
    __syncthreads();
    if (i <= nxg - 1 && j <= nyg - 1) {
      if (i <= nx - 1)
        Fu[(size_t)(i - 2) + (size_t)(nx - 2) * ((size_t)(j - 2) + (size_t)(nyg - 2) * (size_t)(k - 2))] =
            rhs_u_at<false>(g, A, Hu, nu, dPdx, i, j, k);
      if (j <= ny - 1)
        Fv[(size_t)(i - 2) + (size_t)(nxg - 2) * ((size_t)(j - 2) + (size_t)(ny - 2) * (size_t)(k - 2))] =
            rhs_v_at<false>(g, A, Hv, nu, dPdy, i, j, k);
      if (k <= nz - 1)
        Fw[(size_t)(i - 2) + (size_t)(nxg - 2) * ((size_t)(j - 2) + (size_t)(nyg - 2) * (size_t)(k - 2))] =
            rhs_w_at<false>(g, A, Hw, nu, dPdz, i, j, k);
    }
    __syncthreads();  // plane k-1's slot is refilled next
  }
}

// RK3 stage update on the interior of one component
// (src/time_integration.f90:105-107, 134-136, 163-168).
__global__ __launch_bounds__(kB) void k_rk_update(int n1, int n2, int i1, int j1, int k1, int stage,
                                                  double dt, const double* __restrict__ F1,
                                                  const double* __restrict__ F2,
                                                  const double* __restrict__ F3,
                                                  const double* __restrict__ Xo,
                                                  double* __restrict__ X) {
#pragma clang fp contract(off)
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);  // neighbouring tiles on one XCD (shared halo rows)
  const int i = bx * 64 + (threadIdx.x & 63) + 2;
  const int j = by * 4 + (threadIdx.x >> 6) + 2;
  const int k = bz + 2;
  if (i > i1 || j > j1) return;
  const double c11 = 8.0 / 15.0, c21 = 1.0 / 4.0, c22 = 5.0 / 12.0, c31 = 1.0 / 4.0, c32 = 0.0,
               c33 = 3.0 / 4.0;
  const size_t f = (size_t)(i - 2) + (size_t)(i1 - 1) * ((size_t)(j - 2) + (size_t)(j1 - 1) * (size_t)(k - 2));
  const size_t o = fidx(i, j, k, n1, n2);
  double v;
  if (stage == 1) v = Xo[o] + dt * c11 * F1[f];
  else if (stage == 2) v = Xo[o] + dt * (c21 * F1[f] + c22 * F2[f]);
  else v = Xo[o] + dt * (c31 * F1[f] + c32 * F2[f] + c33 * F3[f]);
  X[o] = v;
  (void)k1;
}

// U(2:nx-1, 2:nyg-1, 2:nzg-1) += d  (src/time_integration.f90:190)
__global__ __launch_bounds__(kB) void k_add_u(DevGrid g, double* __restrict__ U, double d) {
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);  // neighbouring tiles on one XCD (shared halo rows)
  const int i = bx * 64 + (threadIdx.x & 63) + 2;
  const int j = by * 4 + (threadIdx.x >> 6) + 2;
  const int k = bz + 2;
  if (i > g.nx - 1 || j > g.nyg - 1) return;
  const size_t o = fidx(i, j, k, g.nx, g.nyg);
  U[o] = U[o] + d;
}

// compute_dt maxima (src/time_integration.f90:228-239): max over the
// centre range of dx/|u| per component, as bit patterns of non-negative
// doubles (+Inf where u = 0).
__global__ __launch_bounds__(kB) void k_dt_max(DevGrid g, const double* __restrict__ U,
                                               const double* __restrict__ V,
                                               const double* __restrict__ W,
                                               unsigned long long* __restrict__ red) {
  __shared__ double sm[3][kB / 64];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 2;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 2;
  double m[3] = {0.0, 0.0, 0.0};
  if (i <= g.nxg - 1 && j <= g.nyg - 1)
    for (int k = blockIdx.z + 2; k <= g.nzg - 1; k += gridDim.z) {
      m[0] = fmax(m[0], g.dxg[i] / fabs(U_(i, j, k)));
      m[1] = fmax(m[1], g.dyg[j] / fabs(V_(i, j, k)));
      m[2] = fmax(m[2], g.dzg[k] / fabs(W_(i, j, k)));
    }
  for (int c = 0; c < 3; c++) {
    double v = m[c];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sm[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = sm[threadIdx.x][0];
    for (int w = 1; w < kB / 64; w++) v = fmax(v, sm[threadIdx.x][w]);
    atomicMax(red + threadIdx.x, (unsigned long long)__double_as_longlong(v));
  }
}

// Bulk-flow partial sums (src/mass_flow.f90:36-44) per (j,k) row, summed
// over i in order: part[2*r] = sum of (H U + H U)*dy/2, part[2*r+1] = norm.
__global__ __launch_bounds__(kB) void k_massflow_rows(DevGrid g, const double* __restrict__ U,
                                                      const double* __restrict__ H,
                                                      double* __restrict__ part) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * kB + threadIdx.x;  // row index over (j = 3..nyg-1, k = 2..nzg-1)
  const int nj = g.nyg - 3, nk = g.nzg - 2;
  if (r >= nj * nk) return;
  const int j = r % nj + 3, k = r / nj + 2;
  double s = 0.0, n = 0.0;
  for (int i = 2; i <= g.nx - 1; i++) {
    const double h1 = H ? H[fidx(i, j, k, g.nx, g.nyg)] : 1.0;
    const double h0 = H ? H[fidx(i, j - 1, k, g.nx, g.nyg)] : 1.0;
    s = s + (h1 * U_(i, j, k) + h0 * U_(i, j - 1, k)) * 0.5 * g.dyg[j - 1];
    n = n + (h1 + h0) * 0.5 * g.dyg[j - 1];
  }
  part[2 * (size_t)r] = s;
  part[2 * (size_t)r + 1] = n;
}

// Ordered sum of the row partials (single workgroup, fixed order).
__global__ __launch_bounds__(kB) void k_sum_pairs(const double* __restrict__ part, int nrow,
                                                  double* __restrict__ out) {
  __shared__ double ss[kB], sn[kB];
  double s = 0.0, n = 0.0;
  for (int r = threadIdx.x; r < nrow; r += kB) {
    s += part[2 * (size_t)r];
    n += part[2 * (size_t)r + 1];
  }
  ss[threadIdx.x] = s;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sn[threadIdx.x] += sn[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ss[0];
    out[1] = sn[0];
  }
}

// Plane sums of compute_statistics on one domain (src/statistics.f90:41-73,
// last-rank branch): one workgroup per row j = 1..nyg.  term_1 is the
// x-centred U (interpolate_x, src/interpolation.f90:21-36), term_2 the
// y-centred V (interpolate_y, :44-58) with term_2(:,1,:) = -term_2(:,2,:)
// and term_2(:,nyg,:) = -term_2(:,nyg-1,:) (src/statistics.f90:47-49).
// sums[q*nyg + j-1] for q = 0 U, 1 V, 2 W, 3 U^2, 4 V^2, 5 W^2, 6 UV over
// i = 2..nxg-2 and k = 2..nzg-2 (W, W^2: k = 2..nz-1); fixed-order sums.
constexpr int kStats = 7;

__global__ __launch_bounds__(kB) void k_plane_sums(DevGrid g, const double* __restrict__ U,
                                                   const double* __restrict__ V,
                                                   const double* __restrict__ W,
                                                   double* __restrict__ sums) {
  __shared__ double sh[kStats][kB];
  const int j = blockIdx.x + 1;
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;
  const int ni = nxg - 3;
  const int nku = g.nzg - 3, nkw = g.nz - 2;
  // term_2 at row j from V rows jv - 1, jv with sign sg (the wall rows mirror)
  const int jv = (j == 1) ? 2 : ((j == nyg) ? nyg - 1 : j);
  const double sg = (j == 1 || j == nyg) ? -1.0 : 1.0;
  double a[kStats] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < ni * nku; t += kB) {
    const int i = t % ni + 2, k = t / ni + 2;
    const double t1 = 0.5 * (U[fidx(i - 1, j, k, nx, nyg)] + U[fidx(i, j, k, nx, nyg)]);
    const double t2 = sg * (0.5 * (V[fidx(i, jv - 1, k, nxg, ny)] + V[fidx(i, jv, k, nxg, ny)]));
    a[0] += t1;
    a[1] += t2;
    a[3] += t1 * t1;
    a[4] += t2 * t2;
    a[6] += t1 * t2;
  }
  for (int t = threadIdx.x; t < ni * nkw; t += kB) {
    const int i = t % ni + 2, k = t / ni + 2;
    const double w = W[fidx(i, j, k, nxg, nyg)];
    a[2] += w;
    a[5] += w * w;
  }
#pragma unroll
  for (int q = 0; q < kStats; q++) sh[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < kStats; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kStats) sums[threadIdx.x * (size_t)nyg + j - 1] = sh[threadIdx.x][0];
}

// output_monitor's field summaries (src/monitor.f90:27-47): per workgroup
// the sums and maxima of U(2:nx-1,2:nyg-1,2:nzg-1), V(2:nxg-1,2:ny-1,2:nzg-1)
// and W(2:nxg-1,2:nyg-1,2:nz-1), and the count of non-finite entries of the
// three (a NaN guard), into part[b*kMon + q]; k_monitor_final reduces the
// kMonBlocks partials in a fixed order.
constexpr int kMon = 7;  // sum U, V, W; max U, V, W; non-finite count
constexpr int kMonBlocks = 256;

__device__ __forceinline__ void mon_field(const double* __restrict__ F, int n1, int n2, int i1, int j1,
                                          int k1, double& sum, double& mx, double& bad) {
  // interior (2..i1, 2..j1, 2..k1) of an (n1, n2, *) array, i fastest
  const size_t ni = i1 - 1, nj = j1 - 1, n = ni * nj * (size_t)(k1 - 1);
  for (size_t t = (size_t)blockIdx.x * kB + threadIdx.x; t < n; t += (size_t)kMonBlocks * kB) {
    const int i = (int)(t % ni) + 2, j = (int)((t / ni) % nj) + 2, k = (int)(t / (ni * nj)) + 2;
    const double v = F[fidx(i, j, k, n1, n2)];
    sum += v;
    mx = fmax(mx, v);
    bad += isfinite(v) ? 0.0 : 1.0;
  }
}

__global__ __launch_bounds__(kB) void k_monitor_part(DevGrid g, const double* __restrict__ U,
                                                     const double* __restrict__ V,
                                                     const double* __restrict__ W,
                                                     double* __restrict__ part) {
  __shared__ double sh[kMon][kB];
  double a[kMon] = {0.0, 0.0, 0.0, -INFINITY, -INFINITY, -INFINITY, 0.0};
  mon_field(U, g.nx, g.nyg, g.nx - 1, g.nyg - 1, g.nzg - 1, a[0], a[3], a[6]);
  mon_field(V, g.nxg, g.ny, g.nxg - 1, g.ny - 1, g.nzg - 1, a[1], a[4], a[6]);
  mon_field(W, g.nxg, g.nyg, g.nxg - 1, g.nyg - 1, g.nz - 1, a[2], a[5], a[6]);
#pragma unroll
  for (int q = 0; q < kMon; q++) sh[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < kMon; q++)
        sh[q][threadIdx.x] = (q >= 3 && q < 6) ? fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + w])
                                                : sh[q][threadIdx.x] + sh[q][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < kMon) part[(size_t)blockIdx.x * kMon + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kB) void k_monitor_final(const double* __restrict__ part,
                                                      double* __restrict__ out) {
  __shared__ double sh[kMon][kB];
#pragma unroll
  for (int q = 0; q < kMon; q++) sh[q][threadIdx.x] = part[(size_t)threadIdx.x * kMon + q];
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < kMon; q++)
        sh[q][threadIdx.x] = (q >= 3 && q < 6) ? fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + w])
                                                : sh[q][threadIdx.x] + sh[q][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < kMon) out[threadIdx.x] = sh[threadIdx.x][0];
}

inline dim3 gpts(int ni, int nj, int nk) { return dim3((ni + 63) / 64, (nj + 3) / 4, nk); }

}  // namespace

// z planes per workgroup of the right-hand sides; C3D_RHS_KC (2, 4, 8, 16)
// overrides the default 8 (tuning knob).
static int rhs_kc() {
  static const int kc = [] {
    const char* e = getenv("C3D_RHS_KC");
    const int v = e ? atoi(e) : 0;
    return (v == 2 || v == 4 || v == 8 || v == 16) ? v : 8;
  }();
  return kc;
}

hipError_t launch_rhs(const DevGrid& g, const double* U, const double* V, const double* W,
                      const double* Hu, const double* Hv, const double* Hw, double nu, double dPdx,
                      double dPdy, double dPdz, double* Fu, double* Fv, double* Fw, hipStream_t s) {
  const int kc = rhs_kc();
  const dim3 grid = gpts(g.nxg - 2, g.nyg - 2, (g.nzg - 2 + kc - 1) / kc);
  switch (kc) {
    case 2: hipLaunchKernelGGL(k_rhs<2>, grid, dim3(kB), 0, s, g, U, V, W, Hu, Hv, Hw, nu, dPdx, dPdy, dPdz, Fu, Fv, Fw); break;
    case 4: hipLaunchKernelGGL(k_rhs<4>, grid, dim3(kB), 0, s, g, U, V, W, Hu, Hv, Hw, nu, dPdx, dPdy, dPdz, Fu, Fv, Fw); break;
    case 16: hipLaunchKernelGGL(k_rhs<16>, grid, dim3(kB), 0, s, g, U, V, W, Hu, Hv, Hw, nu, dPdx, dPdy, dPdz, Fu, Fv, Fw); break;
    default: hipLaunchKernelGGL(k_rhs<8>, grid, dim3(kB), 0, s, g, U, V, W, Hu, Hv, Hw, nu, dPdx, dPdy, dPdz, Fu, Fv, Fw); break;
  }
  return hipGetLastError();
}

hipError_t launch_rk_update(const DevGrid& g, int which, int stage, double dt, const double* F1,
                            const double* F2, const double* F3, const double* Xo, double* X,
                            hipStream_t s) {
  int n1, n2, i1, j1, k1;
  if (which == 0) { n1 = g.nx; n2 = g.nyg; i1 = g.nx - 1; j1 = g.nyg - 1; k1 = g.nzg - 1; }
  else if (which == 1) { n1 = g.nxg; n2 = g.ny; i1 = g.nxg - 1; j1 = g.ny - 1; k1 = g.nzg - 1; }
  else { n1 = g.nxg; n2 = g.nyg; i1 = g.nxg - 1; j1 = g.nyg - 1; k1 = g.nz - 1; }
  hipLaunchKernelGGL(k_rk_update, gpts(i1 - 1, j1 - 1, k1 - 1), dim3(kB), 0, s, n1, n2, i1, j1, k1,
                     stage, dt, F1, F2, F3, Xo, X);
  return hipGetLastError();
}

hipError_t launch_add_u(const DevGrid& g, double* U, double d, hipStream_t s) {
  hipLaunchKernelGGL(k_add_u, gpts(g.nx - 2, g.nyg - 2, g.nzg - 2), dim3(kB), 0, s, g, U, d);
  return hipGetLastError();
}

hipError_t launch_dt_max(const DevGrid& g, const double* U, const double* V, const double* W,
                         unsigned long long* red3, hipStream_t s) {
  hipLaunchKernelGGL(k_dt_max, gpts(g.nxg - 2, g.nyg - 2, g.nzg - 2 < 16 ? g.nzg - 2 : 16), dim3(kB),
                     0, s, g, U, V, W, red3);
  return hipGetLastError();
}

int massflow_rows(const DevGrid& g) { return (g.nyg - 3) * (g.nzg - 2); }

int monitor_part_size() { return kMonBlocks * kMon; }
static_assert(kMonBlocks * kMon == kMonitorPartDoubles, "partials of the monitor kernels");

namespace {
// c3d_monitor's results from the reduced sums r7 (sums U, V, W; maxima U, V,
// W; non-finite count): maxima, sums / vol, count.
__global__ void k_monitor_finish(const double* __restrict__ r7, double vol, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out[0] = r7[3];
  out[1] = r7[4];
  out[2] = r7[5];
  out[3] = r7[0] / vol;
  out[4] = r7[1] / vol;
  out[5] = r7[2] / vol;
  out[6] = r7[6];
}
}  // namespace

hipError_t launch_monitor_finish(const double* r7, double vol, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_monitor_finish, dim3(1), dim3(64), 0, s, r7, vol, out);
  return hipGetLastError();
}

hipError_t launch_monitor(const DevGrid& g, const double* U, const double* V, const double* W,
                          double* part, double* out7, hipStream_t s) {
  static_assert(kMonBlocks == kB, "the final reduction holds one partial per thread");
  hipLaunchKernelGGL(k_monitor_part, dim3(kMonBlocks), dim3(kB), 0, s, g, U, V, W, part);
  hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(kB), 0, s, (const double*)part, out7);
  return hipGetLastError();
}

hipError_t launch_plane_sums(const DevGrid& g, const double* U, const double* V, const double* W,
                             double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_plane_sums, dim3(g.nyg), dim3(kB), 0, s, g, U, V, W, sums);
  return hipGetLastError();
}

hipError_t launch_massflow(const DevGrid& g, const double* U, const double* H, double* part,
                           double* out2, hipStream_t s) {
  const int nrow = massflow_rows(g);
  hipLaunchKernelGGL(k_massflow_rows, dim3((nrow + kB - 1) / kB), dim3(kB), 0, s, g, U, H, part);
  hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kB), 0, s, part, nrow, out2);
  return hipGetLastError();
}

}  // namespace c3d
