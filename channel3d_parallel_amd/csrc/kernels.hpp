// Host-side launchers of the projection kernels (defined in kernels.hip).
#pragma once

#include "c3d_common.hpp"

namespace c3d {

// Grid geometry and solver tables resident on the device.
struct DevGrid {
  int nx, ny, nz;     // face counts
  int nxg, nyg, nzg;  // centre counts incl. ghosts (= n + 1)
  int nxp, nzp, nym;  // FFT lengths and tridiagonal length
  int nxh;            // nxp/2 + 1 (R2C spectrum length)
  // Fortran-indexed spacing tables (entry [i] for 1-based index i):
  const double* dxf;  // x(i) - x(i-1), i = 2..nx
  const double* dyf;  // y(j) - y(j-1), j = 2..ny
  const double* dzf;  // z(k) - z(k-1), k = 2..nz
  const double* dxg;  // xg(i+1) - xg(i), i = 1..nxg-1
  const double* dyg;  // yg(j+1) - yg(j)
  const double* dzg;  // zg(k+1) - zg(k)
  const double* rdxf;  // 1/dxf (same indexing)
  const double* rdyf;  // 1/dyf
  const double* rdzf;  // 1/dzf
  const double* rdxg;  // 1/dxg (same indexing as dxg)
  const double* rdyg;  // 1/dyg
  const double* rdzg;  // 1/dzg
  // nu/spacing factors of the viscous terms (nu of the context)
  const double* nudxf;  // nu/dxf
  const double* nudyf;  // nu/dyf
  const double* nudzf;  // nu/dzf
  const double* nudxg;  // nu/dxg
  const double* nudzg;  // nu/dzg
  const double* kxx;  // [nxp] modified wavenumbers
  const double* kzz;  // [nzp]
  const double* dyy_d;  // [nym]   Dyy(j,j)
  const double* dyy_l;  // [nym-1] Dyy(j+1,j)
  const double* dyy_u;  // [nym-1] Dyy(j,j+1)
  const cpx<double>* twx64;
  const cpx<double>* twz64;
  const cpx<float>* twx32;
  const cpx<float>* twz32;
  // flag word: the velocity-correction kernels (k_grad_correct,
  // k_ib_correct) set it to 1 when a corrected velocity is not finite
  unsigned* nonfinite;
};

// rhs_p offset of Fortran (i,j,k) with bounds (2:nxg-1, 2:nyg-1, 2:nzg).
__host__ __device__ __forceinline__ size_t ridx(int i, int j, int k, int nxg, int nyg) {
  return (size_t)(i - 2) + (size_t)(nxg - 2) * ((size_t)(j - 2) + (size_t)(nyg - 2) * (size_t)(k - 2));
}

// Forward x transform.  from_uvw: input is div(U,V,W) computed on the fly,
// else the rhs_p array.  Writes the half spectrum S[j][kx][kz].
hipError_t launch_fwd_x(const DevGrid& g, bool from_uvw, const double* U, const double* V,
                        const double* W, const double* rhs, cpx<double>* S, hipStream_t s);
hipError_t launch_fft_z(const DevGrid& g, int sign, cpx<double>* S, hipStream_t s);
// Tridiagonal solve along y for every (kx,kz).  nym = 64, 128, 192 use the
// partitioned register solver; other sizes the plain Thomas sweep with CP,
// nym*nxh*nzp doubles of scratch (may be null when tri_part_supported).
bool tri_part_supported(const DevGrid& g);
hipError_t launch_tri(const DevGrid& g, cpx<double>* S, double* CP, hipStream_t s);
// Inverse x transform.  mode 0: out = solution rhs_p (Poisson drop-in).
// mode 1: out = P_interim = solution/dt (projection).  rk1: also write P.
hipError_t launch_inv_x(const DevGrid& g, int mode, const cpx<double>* S, double* out, double* P,
                        double dt, bool rk1, hipStream_t s);
// Refinement pieces (mixed precision: FP32 forward spectra; FP64: the same
// sweeps on FP64 spectra).  launch_fwd_x32: divergence of U, V, W (FP64,
// also stored to b when b != null) -> FP32 spectrum.
hipError_t launch_fwd_x32(const DevGrid& g, const double* U, const double* V, const double* W,
                          double* b, cpx<float>* S, hipStream_t s);
// FP64 spectrum of div(U, V, W), the FP64 divergence also stored to b.
hipError_t launch_fwd_x_b64(const DevGrid& g, const double* U, const double* V, const double* W,
                            double* b, cpx<double>* S, hipStream_t s);
// The refinement residual r = b - M phi formed on the fly as the forward
// transform's input, M = D G on the unique cells plus the spectral solver's
// (0,0)-mode gauge row (gauge: nzp + 1 doubles of scratch).  The FP32 variant
// transforms r rounded to FP32.  max |r| as double bits into rmax[0..64):
// workgroup w maxes into slot w % 64 (zeroed by the caller; the max over the
// slots is max |r|).  rout (may be null) receives the FP64 residual.
hipError_t launch_fwd_x_resid32(const DevGrid& g, const double* b, const double* phi, double* gauge,
                                cpx<float>* S, unsigned long long* rmax, double* rout, hipStream_t s);
hipError_t launch_fwd_x_resid64(const DevGrid& g, const double* b, const double* phi, double* gauge,
                                cpx<double>* S, unsigned long long* rmax, double* rout,
                                hipStream_t s);
hipError_t launch_fft_z32(const DevGrid& g, int sign, cpx<float>* S, hipStream_t s);
// Tridiagonal solve of the FP32 spectrum (tri_part_supported sizes).  low ==
// null: FP64 arithmetic on every mode.  Else FP32 arithmetic on the modes
// with |kxx + kzz| >= lam_thr and FP64 arithmetic on the nlow modes of
// low[] (the kx*nzp + kz indices of exactly the other modes).
hipError_t launch_tri32(const DevGrid& g, cpx<float>* S, const int* low, int nlow, double lam_thr,
                        hipStream_t s);
// The tridiagonal solve of the FP32 spectrum S32 with FP64 arithmetic on
// every mode, the solution stored in FP64 to S64 (tri_part_supported sizes).
hipError_t launch_tri32to64(const DevGrid& g, cpx<float>* S32, cpx<double>* S64, hipStream_t s);
// Inverse x transform with the refinement modes, the potential kept as the
// first solve's phi1 plus a correction phic:
//   0: out = phi, 1: out = phi/dt, 3: out = phi_in + phi,
//   5: out2 = phic' = phi_in2 + phi (phi_in2 may be null: phic' = phi), out
//      unused (launch_grad_correct_split writes P_interim); rk1: P from
//      phi_in + phic';
// rk1 writes P from the same potential.
hipError_t launch_inv_x64_ref(const DevGrid& g, int mode, const cpx<double>* S, const double* phi_in,
                              const double* phi_in2, double* out, double* out2, double* P, double dt,
                              bool rk1, hipStream_t s);
hipError_t launch_inv_x32(const DevGrid& g, int mode, const cpx<float>* S, const double* phi_in,
                          const double* phi_in2, double* out, double* out2, double* P, double dt,
                          bool rk1, hipStream_t s);
// Ghost copies of P (src/poisson.f90:93-98), in the reference order.
hipError_t launch_p_ghosts(const DevGrid& g, double* P, hipStream_t s);
// U -= G(dt * Pint) on the gradient's index ranges.
hipError_t launch_grad_correct(const DevGrid& g, double* U, double* V, double* W, const double* Pint,
                               double dt, hipStream_t s);
// U -= G(phi1) + G(phic), each part differenced on its own; Pint =
// (phi1 + phic)/dt on the whole rhs_p shape.
hipError_t launch_grad_correct_split(const DevGrid& g, double* U, double* V, double* W,
                                     const double* phi1, const double* phic, double* Pint, double dt,
                                     hipStream_t s);
// apply_boundary_conditions on device-resident staggered fields.
hipError_t launch_apply_bc(const DevGrid& g, double* U, double* V, double* W, hipStream_t s);
// Divergence into an rhs_p-shaped array (plane nzg zero).
hipError_t launch_divergence(const DevGrid& g, const double* U, const double* V, const double* W,
                             double* div, hipStream_t s);
// Gradient of an rhs_p-shaped array into zeroed staggered fields.
hipError_t launch_gradient(const DevGrid& g, const double* p, double* Gx, double* Gy, double* Gz,
                           hipStream_t s);
// max |div(U,V,W)| over the rhs_p shape; result as double bits in *red.
hipError_t launch_max_div(const DevGrid& g, const double* U, const double* V, const double* W,
                          unsigned long long* red, hipStream_t s);

// RK3 (rk3.hip): right-hand sides into interior-shaped F arrays; H masks may
// be null (H = 1).
hipError_t launch_rhs(const DevGrid& g, const double* U, const double* V, const double* W,
                      const double* Hu, const double* Hv, const double* Hw, double nu, double dPdx,
                      double dPdy, double dPdz, double* Fu, double* Fv, double* Fw, hipStream_t s);
// Stage update X(int) = Xo(int) + dt*sum a_kl F_l; which 0 U, 1 V, 2 W.
hipError_t launch_rk_update(const DevGrid& g, int which, int stage, double dt, const double* F1,
                            const double* F2, const double* F3, const double* Xo, double* X,
                            hipStream_t s);
hipError_t launch_add_u(const DevGrid& g, double* U, double d, hipStream_t s);
// Fused right-hand side and stage update (rk3_fast.hip), to round-off of the
// reference's operations: reads the stage input U, V, W, writes the updated
// interiors Un, Vn, Wn (a different set of fields) and, for stages 1 and 2,
// the right-hand sides Fu, Fv, Fw; stages 2 and 3 read the step-start
// fields Uo, Vo, Wo and F1.  H masks may be null.
struct RkStage {
  const double *U, *V, *W;
  const double *Hu, *Hv, *Hw;
  double nu, dPdx, dPdy, dPdz, dt;
  const double *Uo, *Vo, *Wo;
  const double *F1u, *F1v, *F1w;
  double *Fu, *Fv, *Fw;
  double *Un, *Vn, *Wn;
};
hipError_t launch_rhs_stage(const DevGrid& g, int stage, const RkStage& a, hipStream_t s);
// compute_dt maxima of dx/|u|, dy/|v|, dz/|w| as double bits into red3[0..2]
// (red3 zeroed by the caller).
hipError_t launch_dt_max(const DevGrid& g, const double* U, const double* V, const double* W,
                         unsigned long long* red3, hipStream_t s);
// Bulk-flow integral and norm of U into out2[0..1]; part holds
// 2*massflow_rows(g) doubles.
int massflow_rows(const DevGrid& g);
hipError_t launch_massflow(const DevGrid& g, const double* U, const double* H, double* part,
                           double* out2, hipStream_t s);
// output_monitor field summaries: out7 = sums of U, V, W over their
// interiors, maxima of the same, and the count of non-finite entries; part
// holds monitor_part_size() doubles of scratch.
int monitor_part_size();
constexpr int kMonitorPartDoubles = 256 * 7;  // monitor_part_size()
// out[0..6] = the results of c3d_monitor (maxima, sums / vol, non-finite
// count) from the reduced sums r7 of launch_monitor, on the device.
hipError_t launch_monitor_finish(const double* r7, double vol, double* out, hipStream_t s);
hipError_t launch_monitor(const DevGrid& g, const double* U, const double* V, const double* W,
                          double* part, double* out7, hipStream_t s);
// Plane sums of compute_statistics over the unique cells of every row
// j = 1..nyg: sums[q*nyg + j-1] for q = 0 x-centred U, 1 y-centred V (wall
// rows mirrored), 2 W, 3 U^2, 4 V^2, 5 W^2, 6 UV (7*nyg doubles).
hipError_t launch_plane_sums(const DevGrid& g, const double* U, const double* V, const double* W,
                             double* sums, hipStream_t s);

// Streaming copy of nbytes (multiple of 16) for bandwidth calibration.
hipError_t launch_copy(const void* src, void* dst, size_t nbytes, hipStream_t s);

}  // namespace c3d
