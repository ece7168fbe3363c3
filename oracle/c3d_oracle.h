/*
 * c3d_oracle.h -- CPU FP64 restatement of the channel3D projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path is tested against this library; it
 * is loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py,
 * never by the product library (channel3d_parallel_amd/), which must fail
 * loudly when its HIP build is missing instead of falling back to this code.
 *
 * Arrays follow the reference's Fortran layout: column-major (x fastest),
 * 1-based Fortran index (i,j,k) stored at 0-based memory offset
 * (i-lo1) + n1*((j-lo2) + n2*(k-lo3)), ghost-inclusive shapes:
 *   U(nx ,nyg,nzg)  V(nxg,ny ,nzg)  W(nxg,nyg,nz )  P(nxg,nyg,nzg)
 *   rhs_p / P_interim (2:nxg-1, 2:nyg-1, 2:nzg)   -- src/initialization.f90:116-155
 * with nxg = nx+1, nyg = ny+1, nzg = nz+1 on a single domain.
 * Complex data is interleaved (re, im) doubles.
 */
#ifndef C3D_ORACLE_H
#define C3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

/* Grid and solver constants (src/input_output.f90:140-247,
 * src/initialization.f90:166-305,346-390). grid_type 0 = uniform y,
 * 1 = tanh-stretched y, 2 = tanh-stretched y with uniform buffers of at
 * least min_buffer_width + 6 dymin at each wall, shifted by +1 (only read
 * for grid_type 2). Returns NULL on invalid sizes. */
orc_grid *orc_grid_new(int nx, int ny, int nz, double Lxp, double Lzp,
                       double Ly, double alpha, int grid_type, double min_buffer_width);
void orc_grid_free(orc_grid *g);
/* what: 0 x[nx] 1 y[ny] 2 z[nz] 3 xg[nxg] 4 yg[nyg] 5 zg[nzg]
 *       6 kxx[nxp] 7 kzz[nzp] 8 Dyy diag[nym] 9 Dyy lower[nym-1]
 *       10 Dyy upper[nym-1]
 *       11 scalars[10] = dx dz dxmin dymin dzmin Lx Ly Lz Lxp Lzp
 *       12 n_uniform[1]
 * Returns the number of doubles written, or -1. */
int orc_grid_get(const orc_grid *g, int what, double *out);

/* Boundary conditions (src/boundary_conditions.f90:20-252). */
void orc_apply_bc(const orc_grid *g, double *U, double *V, double *W);
/* Staggered divergence faces -> centres (src/differential_operators.f90:54-74). */
void orc_divergence(const orc_grid *g, double *div, const double *U,
                    const double *V, const double *W);
/* Staggered gradient centres -> faces (src/differential_operators.f90:16-52). */
void orc_gradient(const orc_grid *g, double *Gx, double *Gy, double *Gz,
                  const double *p);

/* 1-D complex DFT, unnormalised, sign -1 forward / +1 backward
 * (FFTW_FORWARD / FFTW_BACKWARD convention, src/poisson.f90:45,80). */
int orc_fft(int n, int sign, double *data);

/* LAPACK zgtsv restated: complex tridiagonal solve with partial pivoting,
 * one right-hand side (src/poisson.f90:72).  dl[n-1], d[n], du[n-1], b[n]
 * complex interleaved, all overwritten.  Returns INFO. */
int orc_zgtsv(int n, double *dl, double *d, double *du, double *b);

/* solve_poisson_equation (src/poisson.f90:33-101).  rhs has the rhs_p shape
 * and is overwritten with the solution.  If rk_step == 1 and P != NULL the
 * pressure side effect of src/poisson.f90:91-99 is applied.
 * tri_mode 0 = zgtsv restatement, 1 = Thomas without pivoting (diagnostic).
 * nthreads <= 0 uses the OpenMP default. */
void orc_poisson(const orc_grid *g, double *rhs, int rk_step, double dt,
                 double *P, int tri_mode, int nthreads);

/* compute_non_IB_projection (src/projection.f90:22-46) on caller arrays.
 * U,V,W in/out; Uint,Vint,Wint receive the interim copies (may be NULL);
 * Pint receives P_interim (rhs_p shape); rhs is rhs_p scratch/out;
 * P (full shape) receives the rk_step==1 side effect (may be NULL). */
void orc_projection(const orc_grid *g, double *U, double *V, double *W,
                    double *Uint, double *Vint, double *Wint, double *Pint,
                    double *rhs, double *P, int rk_step, double dt,
                    int tri_mode, int nthreads);

/* check_divergence (src/monitor.f90:210-221): max|div| over the rhs_p shape
 * with the extra z plane zero. */
double orc_max_abs_div(const orc_grid *g, const double *U, const double *V,
                       const double *W);

/* ---- RK3 advection-diffusion (c3d_oracle_rk.c) ---- */
/* compute_rhs_u/v/w (src/equations.f90): F in the interior bounds
 * (2:nx-1,2:nyg-1,2:nzg-1) / (2:nxg-1,2:ny-1,2:nzg-1) / (2:nxg-1,2:nyg-1,2:nz-1).
 * H (interior Heaviside mask, field shape) may be NULL for H = 1. */
void orc_rhs_u(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hu, double nu, double dPdx, double *F);
void orc_rhs_v(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hv, double nu, double dPdy, double *F);
void orc_rhs_w(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hw, double nu, double dPdz, double *F);
/* RK3 stage update of one component (which 0 U, 1 V, 2 W), stage 1..3. */
void orc_rk_update(const orc_grid *g, int which, int stage, double dt, const double *F1,
                   const double *F2, const double *F3, const double *Xo, double *Xn);
/* compute_dt (src/time_integration.f90:219-267). */
double orc_compute_dt(const orc_grid *g, const double *U, const double *V, const double *W,
                      double cfl, double nu);
/* H-masked bulk velocity <U> (src/mass_flow.f90:24-53); Hu may be NULL. */
double orc_mean_mass_flow_u(const orc_grid *g, const double *U, const double *Hu);
/* output_monitor field summaries (src/monitor.f90:27-39, 57-59): max U, V,
 * W; sums / (nx-1)(ny-1)(nz-1); count of non-finite entries. */
void orc_monitor(const orc_grid *g, const double *U, const double *V, const double *W,
                 double *out);
/* Plane means of compute_statistics (src/statistics.f90:41-111):
 * out[q*nyg + j-1], q = 0 U, 1 V, 2 W, 3 U^2, 4 V^2, 5 W^2, 6 UV. */
void orc_plane_statistics(const orc_grid *g, const double *U, const double *V, const double *W,
                          double *out);
/* Wall shear of compute_statistics (src/statistics.f90:41-131): utau,
 * wtau, and the plane means Umean, Wmean [nyg] (may be NULL). */
void orc_wall_shear(const orc_grid *g, const double *U, const double *W, double nu,
                    double *utau, double *wtau, double *Umean, double *Wmean);
/* One compute_time_step_RK3 without a body (src/time_integration.f90:86-193);
 * q0 = NAN disables the constant-flow correction. */
void orc_time_step_rk3(const orc_grid *g, double *U, double *V, double *W, double *P,
                       double *Pint, double dt, double nu, double *dPdx, double q0,
                       int nthreads);

/* ---- Immersed boundary (c3d_oracle_ib.c) ----
 * Marker arrays: xb, yb, zb, sb [nb]; normals, ub [3nb] component-major
 * (x block, y block, z block).
 * nzb z slabs of nxb markers; the first nxb1 of each slab are body 1 and the
 * rest body 2 (remove_mean_per_body grouping). */
typedef struct orc_ib orc_ib;
/* deltafnc (src/immersed_boundary_operators.f90:270-294). */
double orc_deltafnc(double x1, double x2, double dr);
/* setup_IB_operators (src/immersed_boundary_operators.f90:19-256); yref
 * [nb] = y_ref_index (NULL: all 1). */
orc_ib *orc_ib_new(const orc_grid *g, int nb, const double *xb, const double *yb,
                   const double *zb, const double *normals, const double *sb,
                   const double *ub, const int *yref, int nzb, int nxb, int nxb1);
/* setup_IB_geometry for the moving walls (src/immersed_boundary_geometry.f90:
 * 176-390): body 1 standing_wave, 2 traveling_wave_x; nb = 2 nxb nzb markers
 * at time t, slab j holding nxb bottom-wall then nxb top-wall markers.
 * p1..p3 = body_param_1..3, ramp = body_ramp_up_time (traveling wave only),
 * Lyc = Ly_channel.  Fills xb, yb, zb, sb [nb], normals, ub [3nb] and
 * yref [nb] (1 bottom, ny top).  Returns the index (1-based) of the first
 * marker whose y lies outside [y(1 + suppy), y(ny - suppy)] (the
 * reference's stop condition), 0 if none, -1 for an invalid body. */
int orc_wave_geometry(const orc_grid *g, int body, int nxb, int nzb, double p1, double p2,
                      double p3, double ramp, double t, double Lyc, double *xb, double *yb,
                      double *zb, double *normals, double *sb, double *ub, int *yref);
void orc_ib_free(orc_ib *ib);
/* stencil s (0 u, 1 v, 2 w, 3 centre): 125 entries per marker, marker-major,
 * Fortran indices of the field (centre: of P). */
void orc_ib_get_stencil(const orc_ib *ib, int s, int *xi, int *yi, int *zi, double *w,
                        double *dxn);
/* E (with_dxn 0) / E1n (with_dxn 1): fields -> out[3nb]. */
void orc_ib_regT(const orc_ib *ib, const double *U, const double *V, const double *W,
                 int with_dxn, double *out);
/* E1n of a centre field in rhs_p shape -> out[nb]. */
void orc_ib_regTc_1n(const orc_ib *ib, const double *P, double *out);
/* R of component comp (0 u, 1 v, 2 w): f[nb] -> full field F. */
void orc_ib_reg(const orc_ib *ib, int comp, const double *f, double *F);
/* remove_mean_per_body (src/immersed_boundary_geometry.f90:400-442), f[nb]. */
void orc_ib_remove_mean(const orc_ib *ib, double *f);
/* compute_heaviside (src/heaviside.f90:20-54); outputs may be NULL. */
void orc_ib_heaviside(orc_ib *ib, double *Hc_ext, double *Hu_ext, double *Hv_ext,
                      double *Hw_ext);
void orc_ib_get_heaviside(const orc_ib *ib, double *E1nHc, double *E1nH, double *Hu_int,
                          double *Hv_int, double *Hw_int);
/* schur (src/projection.f90:98-144), f, Sf [4nb]; rk_step != 1. */
void orc_ib_schur(const orc_ib *ib, const double *f, double *Sf, double dt, double nu,
                  int nthreads);
/* bicgstab (src/projection.f90:146-186); x [4nb] in/out; returns iterations. */
int orc_ib_bicgstab(const orc_ib *ib, double *x, const double *b, double tol, int maxit,
                    double dt, double nu, int rk_step, double *P, double *err, int nthreads);
/* compute_IB_projection (src/projection.f90:48-96) after orc_projection. */
int orc_ib_projection(const orc_ib *ib, double *U, double *V, double *W, const double *Uint,
                      const double *Vint, const double *Wint, double *Pint, double *P,
                      double *fb, double dt, double nu, double tol, int maxit, int rk_step,
                      double *err, int nthreads);
/* check_slip (src/monitor.f90:223-231). */
double orc_ib_max_slip(const orc_ib *ib, const double *U, const double *V, const double *W,
                       const double *fb);

#ifdef __cplusplus
}
#endif
#endif
