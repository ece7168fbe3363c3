// Immersed-boundary operators on the device: delta-kernel stencils,
// interpolation E / E1n, spreading R, per-body mean removal, Schur-complement
// pieces and BiCGSTAB vector kernels (ib.hip).
#pragma once

#include "kernels.hpp"

namespace c3d {
namespace ib {

// Stencil of one staggered location (u, v, w or cell centre) for all
// markers, separable per axis: entry (a, t, l) at [(a*5 + t)*nb + l] for axis
// a = 0 x, 1 y, 2 z and window position t = 0..4.  idx holds the canonical
// Fortran index of the node, phi the 1-D kernel value deltafnc(grid, xb, d),
// dd the signed distance grid - xb.  The 125-point weight of (i, j, k) is
// ((dx*dymin*dz * phi_x[i]) * phi_y[j]) * phi_z[k], the reference's product
// order (src/immersed_boundary_operators.f90:86-91).
struct Stencil {
  int* idx;
  double* phi;
  double* dd;
};

// Marker data and stencils resident on the device.
struct Markers {
  int nb, nzb, nxb, nxb1;
  const double *xb, *yb, *zb, *normals, *sb, *ub;  // normals, ub: [3nb]
  Stencil st[4];  // 0 u, 1 v, 2 w, 3 centre
  double A;       // dx*dymin*dz
  double factor;  // 1/(dx*dymin*dz)
  const int* yref;  // y_ref_index per marker, the face anchoring the y pivots (nullptr: 1)
};

// Node -> entry map of the spreading operator of one component in jagged
// diagonal order: support nodes ranked by descending entry count (node[r]
// its field offset, cnt[r] its entry count), entry i of rank r stored at
// slot[i] + r, so that lanes of consecutive ranks read consecutive
// addresses.  Each node's entries are in ascending marker order.  Per entry:
// q = l*125 + point, the marker ml and the static coefficient
// wfs = (w*factor)*sb[ml] (the reference's product order, so wfs*f[ml]
// rounds as w*factor*sb*f does).
struct Spread {
  int nnodes, nentries;
  const int* node;
  const int* cnt;
  const int* slot;
  const int* q;
  int* ml;
  double* wfs;
};

// Grid data the set-up kernel needs besides DevGrid.
struct SetupGrid {
  const double *x, *y, *z;  // face coordinates, Fortran index i at [i-1]
  double dx, dz, dymin, Lxp, Lzp;
};

// bad (may be nullptr): set to the largest 1-based index of a marker with a
// stencil node off the canonical interior (x in [2, nx-1], z in [2, nz-1],
// y in [2, nyg-1], v: [2, ny-1]); left unchanged when there is none.
hipError_t launch_setup(const DevGrid& g, const SetupGrid& sg, const Markers& m, int* bad,
                        hipStream_t s);

// Moving wall bodies of setup_IB_geometry (src/immersed_boundary_geometry.f90:
// 176-390).  Markers k = i + 2 nxb (j-1) (bottom wall) and k + nxb (top wall),
// i = 1..nxb, j = 1..nzb.  Scalars that depend on t alone are evaluated on the
// host: ct, st = cos, sin(p2 t) (standing wave), amp and damp = amp' (the
// traveling wave's ramp).
struct WaveBody {
  int body;  // 1 standing_wave, 2 traveling_wave_x
  int nxb, nzb, ny;
  double p1, p2, p3;
  double Lxp, Lzp, Lyc, pi;
  double ylo, yhi;  // y(1 + suppy), y(ny - suppy): the reference's stop bounds
};
struct WaveTime {
  double t, ct, st, amp, damp;
};
// xb, yb, zb, sb, yref [nb]; normals, ub [3nb] (z blocks zero).  bad: largest
// 1-based index of a marker outside [ylo, yhi] (unchanged if none).
hipError_t launch_wave_geometry(const WaveBody& w, const WaveTime& wt, double* xb, double* yb,
                                double* zb, double* normals, double* sb, double* ub, int* yref,
                                int* bad, hipStream_t s);

// E (with_dxn false) or E1n (true) of three staggered fields -> out[3nb].
hipError_t launch_interp_uvw(const DevGrid& g, const Markers& m, const double* U, const double* V,
                             const double* W, bool with_dxn, double* out, hipStream_t s);
// E1n of a centre field in rhs_p layout -> out[nb], times sign.
hipError_t launch_interp_c(const DevGrid& g, const Markers& m, const double* P, double sign,
                           double* out, hipStream_t s);
// Schur rows: velocity rows Sf[c*nb+l] = E_c(G(dt*phi) - Reg_c)[l] -
// E1nH[c*nb+l]*dudn[c*nb+l], G the gradient of the rhs_p-shaped phi evaluated
// at the stencil nodes, and the pressure row Sf[3nb+l] = -E1n(phi)[l], in
// one launch.
hipError_t launch_schur_rows(const DevGrid& g, const Markers& m, const double* phi, double dt,
                             const double* Ureg, const double* Vreg, const double* Wreg,
                             const double* E1nH, const double* dudn, double* Sf, hipStream_t s);
// Velocity-rhs rows: out[c*nb+l] = -E_c(U)[l] + ub[c*nb+l].
hipError_t launch_rhs_u(const DevGrid& g, const Markers& m, const double* U, const double* V,
                        const double* W, double* out, hipStream_t s);
// aux[c*nb+l] = dt*nu*f[c*nb+l] - dt*f[3nb+l]*n[c*nb+l]   (c = 0..2)
hipError_t launch_aux(const Markers& m, const double* f, double dt, double nu, double* aux,
                      hipStream_t s);
// R: Reg_c[node] = sum over entries of w*factor*sb*f_c, written at support
// nodes only (other nodes keep the zeros they were created with).
hipError_t launch_spread(const Markers& m, const Spread& sp, int comp, const double* f,
                         double* F, hipStream_t s);
// The three components at once: F_c from f + c*nb with map sp[c].
hipError_t launch_spread3(const Markers& m, const Spread sp[3], const double* f, double* F0,
                          double* F1, double* F2, hipStream_t s);
// Fills sp.ml and sp.wfs of component comp from the stencils (once per body).
hipError_t launch_spread_coeffs(const Markers& m, const Spread& sp, int comp, hipStream_t s);

// Device construction of a Spread map from the stencil indices (ibmap.hip),
// identical to the host construction: entries sorted by node offset (stable,
// so ascending l*125 + point within a node), ranks by descending count
// (stable in the offset).  The Spread arrays node, cnt, q, ml, wfs hold
// nb*125 entries and slot nb*125 + 1; work is the scratch of
// spread_map_work_bytes(nb, nf).  The node count is written to *nnodes
// (device int); the caller copies it into sp.nnodes before spreading.
size_t spread_map_work_bytes(int nb, size_t nf);
hipError_t build_spread_map(const Markers& m, int comp, int n1, int n2, size_t nf, void* work,
                            Spread& sp, int* nnodes, hipStream_t s);
// f[nb] -= per-body mean; then f -= coef*pj when coef != nullptr.  part
// holds 128 doubles of scratch.
hipError_t launch_remove_mean(const Markers& m, double* f, const double* coef, const double* pj,
                              double* part, hipStream_t s);
// U = (Uint - Ureg) - G(dt*(Pint - phi)) on the gradient's index ranges,
// (Uint - Ureg) elsewhere (src/projection.f90:83-92).
hipError_t launch_ib_correct(const DevGrid& g, const double* Uint, const double* Vint,
                             const double* Wint, const double* Ureg, const double* Vreg,
                             const double* Wreg, const double* Pint, const double* phi, double dt,
                             double* U, double* V, double* W, hipStream_t s);

// Element-wise helpers.
// y = a + b*x over n entries
hipError_t launch_affine(const double* x, double a, double b, double* y, size_t n, hipStream_t s);
// Heaviside face interpolation of the rhs_p-shaped Hc into zeroed fields
// (src/heaviside.f90:40-42).
hipError_t launch_h_interp(const DevGrid& g, const double* Hc, double* Hu, double* Hv, double* Hw,
                           hipStream_t s);

// BiCGSTAB (src/projection.f90:146-186) device scalars and vector kernels.
// Scalar slots: 0 rho_o, 1 alpha, 2 om, 3 rho_n, 4 bta, 5 err.
constexpr int kDotBlocks = 128;
// part[slot*kDotBlocks + b] = block b's partial of sum a*b (slot selects one of
// the partial arrays); two dots at once when a2 != nullptr (slot + 1).
hipError_t launch_dot(const double* a, const double* b, const double* a2, const double* b2,
                      size_t n, double* part, int slot, hipStream_t s);
// Scalar steps: 0 initialise, 4 err (also to *host_err when not null), 5
// out[0] = 1/rms over n_r entries of the squares summed in slot sr, out[1]
// likewise for slot sf (0 where the sum is 0).
hipError_t launch_norm_scales(const double* part, int sr, size_t n_r, int sf, size_t n_f,
                              double* out, hipStream_t s);
// minimal-residual step sc[6] = <r,Sd>/<Sd,Sd> from partial slots 2 and 3.
hipError_t launch_bicg_scalar(int step, double* sc, const double* part, double* host_err,
                              hipStream_t s);
// rho_n = rhat.r (slot 0), bta = (rho_n/rho_o)*(alpha/om) -> sc[3], sc[4];
// p = r + bta*(p - om*v)
hipError_t launch_bicg_p(const double* r, const double* v, double* sc, const double* part, double* p,
                         size_t n, hipStream_t s);
// alpha = rho_n / rhat.v (slot 1) -> sc[1], rho_o = rho_n -> sc[0];
// h = x + alpha*p;  s = r - alpha*v
hipError_t launch_bicg_hs(const double* x, const double* p, const double* r, const double* v,
                          double* sc, const double* part, double* h, double* sv, size_t n,
                          hipStream_t s);
// om = t.s / t.t (slots 2, 3) -> sc[2]; x = h + om*s;  r = s - om*t;
// partials of r.r into slot 4 and of rhat.r into slot 0
hipError_t launch_bicg_xr(const double* h, const double* sv, const double* t, const double* rhat,
                          double* sc, double* part, double* x, double* r, size_t n, hipStream_t s);
// r = b - r (b == null: r unchanged); rhat = r; p = v = 0
hipError_t launch_bicg_init(const double* b, double* r, double* rhat, double* p, double* v,
                            size_t n, hipStream_t s);
// x += a d; r -= a Sd with a = sc[6]
hipError_t launch_line_step(double* x, const double* d, double* r, const double* Sd,
                            const double* sc, size_t n, hipStream_t s);
// History-projection start for successive right-hand sides of one Schur
// operator: up to kHist earlier (force, rhs) pairs are orthonormalised in rhs
// space (Gram-Schmidt, two passes) with the same combinations applied to the
// forces, and the start is sum_j <Q_j, b> X_j.
constexpr int kHist = 8;
constexpr double kHistDrop = 1e-12;  // relative norm below which a column is dependent
struct VecSet {
  const double* v[kHist + 1];
  int k;
};
// part[j*kDotBlocks + b] = block b's partial of <A.v[j], q>, j < A.k (<= kHist+1)
hipError_t launch_mdot(const VecSet& A, const double* q, size_t n, double* part, hipStream_t s);
// coef[j] = fixed-order sum of the partials of slot j, j < k
hipError_t launch_mreduce(const double* part, int k, double* coef, hipStream_t s);
// qa = (set ? 0 : qa) + sgn*sum_j coef[j] A.v[j];  xb likewise with B.v[j]
// (either output may be null; B.k must equal A.k when xb is given)
hipError_t launch_combine(const VecSet& A, const VecSet& B, const double* coef, double sgn, bool set,
                          double* qa, double* xb, size_t n, hipStream_t s);
// q, x /= sqrt(*nrm2), or q = x = 0 when *nrm2 <= kHistDrop^2 * *raw2
hipError_t launch_hscale(double* q, double* x, const double* nrm2, const double* raw2, size_t n,
                         hipStream_t s);
// max |E u - E1nH*dudn - ub| as double bits into *red (zeroed by the caller).
hipError_t launch_max_slip(const DevGrid& g, const Markers& m, const double* U, const double* V,
                           const double* W, const double* E1nH, const double* fb,
                           unsigned long long* red, hipStream_t s);

}  // namespace ib
}  // namespace c3d
