// Fused right-hand side and stage update of the RK3 time step
// (src/equations.f90:24-423, src/time_integration.f90:101-107, 130-136,
// 159-168) for gfx950, FP64, to round-off tolerance of the reference's
// operations: the spacing quotients are products with tabulated reciprocals
// and FMA contraction is on (rk3.hip keeps the bitwise form).
//
// One kernel per stage reads the stage's velocity through the LDS tile of
// k_rhs, evaluates F_s at each point and writes, in the same pass,
// X_s = X_0 + dt sum_l a_sl F_l into a second set of fields (the stage's
// input is still read by the neighbouring workgroups' halos, so the update
// cannot be in place).  F_s is stored for the later stages that use it.
#include <cstdlib>

#include "kernels.hpp"
#include "rk3_point.hpp"

namespace c3d {

namespace {

constexpr int kBf = 256;

// Wray coefficients as the reference writes them (src/time_integration.f90:
// 105-107, 134-136, 163-168): stage 2 uses F1, F2; stage 3 uses F1 and F3
// (the F2 coefficient of stage 3 is zero).
constexpr double kC11 = 8.0 / 15.0, kC21 = 1.0 / 4.0, kC22 = 5.0 / 12.0, kC31 = 1.0 / 4.0,
                 kC33 = 3.0 / 4.0;

// PIPE: the loads of plane k+2 are issued before plane k is evaluated
// (software pipelining through registers); otherwise each plane is loaded
// into LDS right before the barrier that precedes its use.
template <int KC, int STAGE, bool PIPE>
__device__ __forceinline__ void rhs_stage_body(const DevGrid& g, const RkStage& a) {
  __shared__ double tile[3 * 3 * kTP];
  __shared__ double xt[kXT * kTX];
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);
  const int i0 = bx * 64 + 2, j0 = by * 4 + 2, kb = bz * KC + 2;
  // j is the same on all lanes of a wavefront (64 x positions of one row):
  // wave-uniform, so its spacing factors are scalar loads
  const int i = i0 + (threadIdx.x & 63);
  const int j = __builtin_amdgcn_readfirstlane(j0 + (int)(threadIdx.x >> 6));
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  // the tile's x spacing factors, i0-1 .. i0+64, clamped to the tables'
  // entries 0 .. nx+1 (lanes past the domain read but do not use them);
  // made visible by the first plane's barrier below
  for (int e = threadIdx.x; e < kXT * kTX; e += kBf) {
    const int t = e / kTX, ii = min(i0 - 1 + e % kTX, nx + 1);
    const double* src = t == kXrdxg ? g.rdxg : t == kXrdxf ? g.rdxf : t == kXnudxf ? g.nudxf : g.nudxg;
    xt[e] = src[ii];
  }
  TileAccX A{{tile, i0, j0, kb}, xt};
  // planes kb-1 and kb in LDS, kb+1 in registers; with PIPE, at plane k the
  // loads of plane k+2 are in flight while k is evaluated
  PlaneRegs nxt;
  load_plane(g, a.U, a.V, a.W, A, kb - 1, nxt);
  store_plane(tile, A, kb - 1, nxt);
  load_plane(g, a.U, a.V, a.W, A, kb, nxt);
  store_plane(tile, A, kb, nxt);
  if (PIPE) load_plane(g, a.U, a.V, a.W, A, kb + 1, nxt);
  const double dt = a.dt;
  for (int kk = 0; kk < KC; kk++) {
    const int k = kb + kk;
    if (k > nzg - 1) break;  // uniform across the workgroup
    if (!PIPE) load_plane(g, a.U, a.V, a.W, A, k + 1, nxt);
    store_plane(tile, A, k + 1, nxt);  // its slot held plane k-2, done with last iteration
    __syncthreads();
    if (PIPE && kk + 1 < KC && k + 1 <= nzg - 1) load_plane(g, a.U, a.V, a.W, A, k + 2, nxt);
    if (i <= nxg - 1 && j <= nyg - 1) {
      if (i <= nx - 1) {  // U(2:nx-1, 2:nyg-1, 2:nzg-1)
        const size_t f = (size_t)(i - 2) + (size_t)(nx - 2) * ((size_t)(j - 2) + (size_t)(nyg - 2) * (size_t)(k - 2));
        const size_t o = fidx(i, j, k, nx, nyg);
        const double F = rhs_u_at<true>(g, A, a.Hu, a.nu, a.dPdx, i, j, k);
        double x;
        if constexpr (STAGE == 1) x = A.U(i, j, k) + (dt * kC11) * F;
        else if constexpr (STAGE == 2) x = a.Uo[o] + dt * (kC21 * a.F1u[f] + kC22 * F);
        else x = a.Uo[o] + dt * (kC31 * a.F1u[f] + kC33 * F);
        if constexpr (STAGE < 3) a.Fu[f] = F;
        a.Un[o] = x;
      }
      if (j <= ny - 1) {  // V(2:nxg-1, 2:ny-1, 2:nzg-1)
        const size_t f = (size_t)(i - 2) + (size_t)(nxg - 2) * ((size_t)(j - 2) + (size_t)(ny - 2) * (size_t)(k - 2));
        const size_t o = fidx(i, j, k, nxg, ny);
        const double F = rhs_v_at<true>(g, A, a.Hv, a.nu, a.dPdy, i, j, k);
        double x;
        if constexpr (STAGE == 1) x = A.V(i, j, k) + (dt * kC11) * F;
        else if constexpr (STAGE == 2) x = a.Vo[o] + dt * (kC21 * a.F1v[f] + kC22 * F);
        else x = a.Vo[o] + dt * (kC31 * a.F1v[f] + kC33 * F);
        if constexpr (STAGE < 3) a.Fv[f] = F;
        a.Vn[o] = x;
      }
      if (k <= nz - 1) {  // W(2:nxg-1, 2:nyg-1, 2:nz-1)
        const size_t f = (size_t)(i - 2) + (size_t)(nxg - 2) * ((size_t)(j - 2) + (size_t)(nyg - 2) * (size_t)(k - 2));
        const size_t o = fidx(i, j, k, nxg, nyg);
        const double F = rhs_w_at<true>(g, A, a.Hw, a.nu, a.dPdz, i, j, k);
        double x;
        if constexpr (STAGE == 1) x = A.W(i, j, k) + (dt * kC11) * F;
        else if constexpr (STAGE == 2) x = a.Wo[o] + dt * (kC21 * a.F1w[f] + kC22 * F);
        else x = a.Wo[o] + dt * (kC31 * a.F1w[f] + kC33 * F);
        if constexpr (STAGE < 3) a.Fw[f] = F;
        a.Wn[o] = x;
      }
    }
    __syncthreads();  // plane k-1's slot is refilled next
  }
}

template <int KC, int STAGE, bool PIPE>
__global__ __launch_bounds__(kBf) void k_rhs_stage(DevGrid g, RkStage a) {
  rhs_stage_body<KC, STAGE, PIPE>(g, a);
}

// the pipelined form held to 128 VGPRs (4 waves per SIMD instead of 3, a
// few spilled registers)
template <int KC, int STAGE>
__global__ __launch_bounds__(kBf) __attribute__((amdgpu_waves_per_eu(4))) void k_rhs_stage_w4(
    DevGrid g, RkStage a) {
  rhs_stage_body<KC, STAGE, true>(g, a);
}

inline dim3 gpts_f(int ni, int nj, int nk) { return dim3((ni + 63) / 64, (nj + 3) / 4, nk); }

// Plane loading of the fused kernel: 0 load-then-barrier (the default), 1
// pipelined, 2 pipelined at 4 waves per SIMD.  Measured at 384x193x384
// before the x factors moved to LDS: 5.43, 5.61 (3 instead of 4 waves per
// SIMD) and 5.72 ms (spills) per RK3 step.
// C3D_RHS_PIPE overrides the default (tuning knob).
int rhs_pipe() {
  static const int v = [] {
    const char* e = getenv("C3D_RHS_PIPE");
    const int x = e ? atoi(e) : 0;
    return (x >= 0 && x <= 2) ? x : 0;
  }();
  return v;
}

template <int STAGE>
void launch_stage_kernel(const DevGrid& g, const RkStage& a, dim3 grid, hipStream_t s) {
  constexpr int KC = 8;
  switch (rhs_pipe()) {
    case 1: hipLaunchKernelGGL((k_rhs_stage<KC, STAGE, true>), grid, dim3(kBf), 0, s, g, a); break;
    case 2: hipLaunchKernelGGL((k_rhs_stage_w4<KC, STAGE>), grid, dim3(kBf), 0, s, g, a); break;
    default: hipLaunchKernelGGL((k_rhs_stage<KC, STAGE, false>), grid, dim3(kBf), 0, s, g, a); break;
  }
}

}  // namespace

hipError_t launch_rhs_stage(const DevGrid& g, int stage, const RkStage& a, hipStream_t s) {
  constexpr int KC = 8;
  const dim3 grid = gpts_f(g.nxg - 2, g.nyg - 2, (g.nzg - 2 + KC - 1) / KC);
  switch (stage) {
    case 1: launch_stage_kernel<1>(g, a, grid, s); break;
    case 2: launch_stage_kernel<2>(g, a, grid, s); break;
    case 3: launch_stage_kernel<3>(g, a, grid, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace c3d
