/*
 * c3d.h -- C ABI of the MI355X (gfx950) channel3D projection library.
 *
 * Drop-in boundary for the fractional-step projection path of the
 * reference's Fortran modules (poisson, projection, differential_operators,
 * boundary_conditions, monitor).  The reference exposes module procedures
 * that act on module-global state; here one opaque context (c3d_ctx) holds
 * that state on one GPU, and a Fortran host binds these entry points with
 * iso_c_binding (see the bind(C) interfaces in
 * channel3d_parallel_amd/fortran/c3d_bindings.f90).
 *
 * Conventions
 *  - Plain pointers and sizes, no framework types.  Host arrays use the
 *    reference's Fortran layout: column-major (x fastest), ghost-inclusive,
 *    FP64:  U(nx,nyg,nzg)  V(nxg,ny,nzg)  W(nxg,nyg,nz)  P(nxg,nyg,nzg)
 *    rhs_p / P_interim with bounds (2:nxg-1, 2:nyg-1, 2:nzg)
 *    where nxg = nx+1, nyg = ny+1, nzg = nz+1 and nx,ny,nz are the face
 *    counts (nx_global, ny_global, nz_global of the reference namelist).
 *  - Every call returns 0 on success or a negative C3D_E* code; the message
 *    is available from c3d_last_error().  Calls are synchronous on return
 *    (one HIP stream per context).  One host thread per context.
 *  - Fields stay resident in device memory between calls.
 */
#ifndef C3D_H
#define C3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define C3D_OK 0
#define C3D_EINVAL (-1)   /* invalid argument or size */
#define C3D_EHIP (-2)     /* HIP runtime error */
#define C3D_ENOMEM (-3)   /* device allocation failed */
#define C3D_ENOCONV (-4)  /* iterative solve hit its iteration limit */
#define C3D_ENAN (-5)     /* non-finite value detected */

#define C3D_PREC_FP64 0  /* plain FP64 path */
#define C3D_PREC_MIXED 1 /* FP32 forward transforms + FP64 refinement;
                            wall-normal length ny-1 in {64, 128, 192} */

typedef struct c3d_ctx c3d_ctx;

/* Grid and physics description.  Mirrors the &params namelist entries of
 * src/input_output.f90:27-40 that the projection path reads. */
typedef struct c3d_desc {
  int nx, ny, nz;   /* nx_global, ny_global, nz_global (face counts, even nx,nz) */
  int grid_type;    /* 0 uniform y, 1 tanh-stretched y, 2 tanh-stretched y with
                       uniform wall buffers (src/input_output.f90:156-247) */
  double Lxp;       /* periodic length in x (default 2*pi) */
  double Lzp;       /* periodic length in z (default pi) */
  double Ly;        /* Ly_channel (default 2) */
  double alpha;     /* alpha_stretch */
  double nu;        /* kinematic viscosity */
  int precision;    /* C3D_PREC_FP64 or C3D_PREC_MIXED */
  double min_buffer_width; /* grid_type 2: each buffer holds at least this
                              plus 6 dymin (src/input_output.f90:195) */
} c3d_desc;

/* Creates a context on the current HIP device: builds the grid, spacing
 * tables, modified wavenumbers and Dyy (src/initialization.f90:166-390),
 * allocates the resident fields.  Replaces `initialize`
 * (src/initialization.f90:24-460) for the projection state. */
int c3d_create(const c3d_desc *desc, c3d_ctx **out);
int c3d_destroy(c3d_ctx *ctx);
const char *c3d_last_error(const c3d_ctx *ctx);

/* Copies a grid vector to host memory (n = capacity in doubles).
 * what: 0 x 1 y 2 z 3 xg 4 yg 5 zg 6 kxx 7 kzz 8 Dyy diag 9 Dyy lower
 *       10 Dyy upper 11 scalars{dx,dz,dxmin,dymin,dzmin,Lx,Ly,Lz,Lxp,Lzp}
 *       12 {n_uniform} (buffer points per wall, grid_type 2; else 1).
 * Returns the number of values written (>= 0) or an error code. */
int c3d_grid_get(const c3d_ctx *ctx, int what, double *out, int n);

/* Host <-> device copies of the resident state (reference shapes above).
 * NULL pointers are skipped.  P_interim uses the rhs_p shape. */
int c3d_set_fields(c3d_ctx *ctx, const double *U, const double *V, const double *W,
                   const double *P);
int c3d_get_fields(c3d_ctx *ctx, double *U, double *V, double *W, double *P, double *P_interim);

/* apply_boundary_conditions(U,V,W) on the resident velocity
 * (src/boundary_conditions.f90:20-55, single domain). */
int c3d_apply_boundary_conditions(c3d_ctx *ctx);

/* divergence(div, U, V, W) of the resident velocity into a host array of
 * the rhs_p shape (src/differential_operators.f90:54-74). */
int c3d_divergence(c3d_ctx *ctx, double *div);

/* gradient(Gpx, Gpy, Gpz, p) of a host rhs_p-shaped array into host
 * staggered arrays (src/differential_operators.f90:16-52). */
int c3d_gradient(c3d_ctx *ctx, const double *p, double *Gpx, double *Gpy, double *Gpz);

/* Host-array forms of the procedures above, for callers that pass their own
 * arrays (the reference calls apply_boundary_conditions and divergence on
 * U_reg/V_reg/W_reg and the Heaviside fields, src/projection.f90:73,76,116,119,
 * src/heaviside.f90:33,35,43,49).  The arrays are staged in device scratch
 * owned by the context (allocated by c3d_create); the resident state is not
 * touched and nothing is allocated per call. */
/* apply_boundary_conditions(U_, V_, W_) on host arrays in place
 * (src/boundary_conditions.f90:20-55). */
int c3d_apply_boundary_conditions_arrays(c3d_ctx *ctx, double *U, double *V, double *W);
/* divergence(div_, U_, V_, W_) of host arrays into a host rhs_p-shaped array
 * (src/differential_operators.f90:54-74). */
int c3d_divergence_arrays(c3d_ctx *ctx, const double *U, const double *V, const double *W,
                          double *div);

/* solve_poisson_equation(rhs) on a host rhs_p-shaped array, in place
 * (src/poisson.f90:33-101).  If rk_step == 1 the resident P receives
 * rhs/(dt*8/15) with neighbour-copied ghosts, as at src/poisson.f90:91-99. */
int c3d_solve_poisson_equation(c3d_ctx *ctx, double *rhs, int rk_step, double dt);

/* compute_non_IB_projection on the resident fields
 * (src/projection.f90:22-46): U <- U - G(phi), P_interim <- phi/dt, and
 * the P side effect at rk_step == 1.  Boundary faces are left as the
 * reference leaves them; callers apply c3d_apply_boundary_conditions after,
 * as src/time_integration.f90:117-123 does. */
int c3d_compute_non_ib_projection(c3d_ctx *ctx, double dt, int rk_step);

/* check_divergence: max|div(U,V,W)| over the rhs_p shape
 * (src/monitor.f90:210-221). */
int c3d_check_divergence(c3d_ctx *ctx, double *max_div);

/* Wall shear of compute_statistics on the resident U and W
 * (src/statistics.f90:41-131, one domain): plane means of the x-centred U
 * and of W over the unique cells, wall gradients from rows 1, 2 and nyg-1,
 * nyg, zero wall Reynolds stresses, utau = sqrt((nu*dUdy_b - nu*dUdy_t)/Ly)
 * and wtau likewise (NaN when the argument is negative, as the reference's
 * **0.5d0).  Umean, Wmean [nyg] receive the plane means when not NULL. */
int c3d_wall_shear(c3d_ctx *ctx, double *utau, double *wtau, double *Umean, double *Wmean);
/* Field summaries of output_monitor on the resident fields
 * (src/monitor.f90:27-39, 57-59): out[0..2] = maxval of U, V, W over their
 * interiors (U(2:nx-1,2:nyg-1,2:nzg-1), V(2:nxg-1,2:ny-1,2:nzg-1),
 * W(2:nxg-1,2:nyg-1,2:nz-1)); out[3..5] = their sums divided by
 * (nx-1)*(ny-1)*(nz-1), as the reference prints "Mean U/V/W"; out[6] = the
 * number of non-finite entries in those interiors (a NaN guard for a host
 * loop; the reference has none). */
int c3d_monitor(c3d_ctx *ctx, double *out);
/* The results of c3d_monitor for device arrays U(nx,nyg,nzg),
 * V(nxg,ny,nzg), W(nxg,nyg,nz) that the caller owns, into the device array
 * out[7], enqueued on `stream` (a hipStream_t; NULL: the null stream) and
 * not waited for.  No context and no allocation: the kernels' scratch (the
 * 256 x 7 partials and the reduced sums, 14 KB whatever the grid) is a
 * static device array of the library, one per GPU, held until the process
 * ends.  Face counts 3 .. 1000000 each; C3D_EINVAL otherwise or for NULL
 * arrays.  Calls that overlap on one GPU share that scratch and race: serialise
 * them (one stream, or events).  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc). */
int c3d_monitor_noalloc(int nx, int ny, int nz, const double *U, const double *V, const double *W,
                        double *out, void *stream);
/* Plane-mean profiles of compute_statistics on the resident fields
 * (src/statistics.f90:41-111, one domain): means[q*nyg + j-1] for q = 0
 * Umean, 1 Vmean, 2 Wmean, 3 U2mean, 4 V2mean, 5 W2mean, 6 UVmean, from the
 * x-centred U and the y-centred V (wall rows mirrored) over the unique
 * cells; 7*nyg doubles. */
int c3d_plane_statistics(c3d_ctx *ctx, double *means);

/* ---- RK3 time integration ------------------------------------------------ */

/* Mean pressure gradients dPdx, dPdy, dPdz used by the right-hand sides
 * (module variables of src/global.f90:176 in the reference). */
int c3d_set_pressure_gradient(c3d_ctx *ctx, double dPdx, double dPdy, double dPdz);
int c3d_get_pressure_gradient(const c3d_ctx *ctx, double *dPdx, double *dPdy, double *dPdz);

/* Uo = U, Vo = V, Wo = W (src/time_integration.f90:92-94). */
int c3d_save_state(c3d_ctx *ctx);

/* compute_rhs_u/v/w(U,V,W, F) of the resident velocity into the stage slot
 * (1..3) Fu<slot>, Fv<slot>, Fw<slot> (src/equations.f90:24-423). */
int c3d_compute_rhs(c3d_ctx *ctx, int slot);
/* Copies stage slot right-hand sides to host arrays with the interior
 * shapes (nx-2,nyg-2,nzg-2), (nxg-2,ny-2,nzg-2), (nxg-2,nyg-2,nz-2). */
int c3d_get_rhs(c3d_ctx *ctx, int slot, double *Fu, double *Fv, double *Fw);

/* compute_rhs_u/v/w(U_, V_, W_, rhs) of host arrays (src/equations.f90:24-29,
 * 160-165, 295-300) into host arrays of the interior shapes of c3d_get_rhs;
 * NULL outputs are skipped.  Uses the pressure gradients of
 * c3d_set_pressure_gradient and the body's H masks as the resident form does;
 * the resident state is not touched. */
int c3d_compute_rhs_arrays(c3d_ctx *ctx, const double *U, const double *V, const double *W,
                           double *Fu, double *Fv, double *Fw);

/* Stage update U(int) = Uo(int) + dt*sum_l rk_coef(rk_step,l)*F_l with the
 * Wray coefficients (src/time_integration.f90:105-107,134-136,163-168). */
int c3d_rk3_update(c3d_ctx *ctx, int rk_step, double dt);

/* compute_dt (src/time_integration.f90:219-267), including the reference's
 * maximum of dx/|u| and the 0.1 cap; cfl < 0 gives dt = -cfl. */
int c3d_compute_dt(c3d_ctx *ctx, double cfl, double *dt);

/* compute_mean_mass_flow_U: H-masked trapezoid bulk velocity of the
 * resident U (src/mass_flow.f90:24-53). */
int c3d_mean_mass_flow_u(c3d_ctx *ctx, double *qflow);
/* compute_mean_mass_flow_U(U, Qflow) of a host array U(nx,nyg,nzg)
 * (src/mass_flow.f90:24-53); compute_dPx_for_constant_mass_flow(U, dPdx)
 * (:144-175) is Qflow_x_0 minus this value. */
int c3d_mean_mass_flow_u_array(c3d_ctx *ctx, const double *U, double *qflow);

/* U(2:nx-1,2:nyg-1,2:nzg-1) += d (src/time_integration.f90:190). */
int c3d_add_to_u(c3d_ctx *ctx, double d);

/* One compute_time_step_RK3 (src/time_integration.f90:86-193) on the
 * resident state: three stages of right-hand sides, stage update,
 * boundary conditions, non-IB projection, boundary conditions (with a body
 * set up: then the IB projection and boundary conditions, :118-123); then, if
 * x_mass_cte == 1, the constant-flow correction toward q0 (Qflow_x_0),
 * which also updates dPdx.  Returns C3D_ENAN when a velocity written by the
 * step's projections is not finite (the NaN stop of compute_statistics,
 * src/statistics.f90:142-144, checked at every step). */
int c3d_time_step_rk3(c3d_ctx *ctx, double dt, int x_mass_cte, double q0);

/* Right-hand sides of c3d_time_step_rk3.  on == 0 (the default, unless the
 * environment sets C3D_RHS_EXACT=1): each stage evaluates compute_rhs_u/v/w
 * and the stage update in one fused pass, with the spacing quotients as
 * products with tabulated reciprocals and FMA contraction -- the reference's
 * operations to round-off (relative 1e-12 after one step of the 32x64x32
 * channel).  on != 0: the reference's divisions, no contraction, separate
 * update pass: bitwise the oracle, as c3d_compute_rhs and c3d_rk3_update
 * always are.  In the fused mode the stage slots hold F1 and F2 of the last
 * step afterwards (F3 is not stored).  No reference counterpart: a choice
 * of arithmetic for src/equations.f90:24-423 and
 * src/time_integration.f90:101-168. */
int c3d_set_rhs_exact(c3d_ctx *ctx, int on);

/* Timing of resident projections with HIP events on the context stream:
 * a pass of `iters` back-to-back projections with an event between kernel
 * stages gives kernel_ms[0..nk), the average ms per stage (order from
 * c3d_stage_name, nk <= c3d_stage_count()); a second pass of `iters`
 * projections between two events only gives total_ms, the average ms per
 * projection.  No reference counterpart: measurement hook for the hot path. */
int c3d_time_projection(c3d_ctx *ctx, int iters, double dt, double *total_ms, double *kernel_ms,
                        int nk);
int c3d_stage_count(void);
/* Refinement sweeps per projection (0..8): C3D_PREC_MIXED contexts default
 * to 1, C3D_PREC_FP64 contexts to 0 (the reference's single solve).  A first
 * solve gives the potential phi1; each sweep transforms the FP64 residual
 * b - M(phi1 + phic) of the projection's own staggered operator (M = D G plus
 * the spectral solver's (0,0)-mode gauge row) and adds its solution to the
 * correction phic; the velocity is corrected by G(phi1) + G(phic), the two
 * parts differenced separately, and P_interim = (phi1 + phic)/dt.  Mixed
 * passes run the forward x/z transforms on FP32 spectra and the tridiagonal
 * solution, inverse z and x transforms in FP64 (FP32 throughout if the
 * environment sets C3D_MIXED_INV64=0 before c3d_create; the default sweep
 * count is then 2).  On an FP64 context the sweeps are FP64 defect
 * corrections.  The first call asking for sweeps allocates their device
 * buffers (3-4 rhs_p-sized FP64 arrays); C3D_ENOMEM if that fails.  No
 * reference counterpart: the reference solves once in FP64
 * (src/poisson.f90:33-101). */
int c3d_set_refinement(c3d_ctx *ctx, int nref);
/* max |residual| entering each sweep of the last projection,
 * rmax[0..min(n, nref)); returns the count written. */
int c3d_refinement_residuals(c3d_ctx *ctx, double *rmax, int n);
/* Streaming device-to-device copy of nbytes (multiple of 16), `iters`
 * times; *ms = average ms per copy.  No reference counterpart: calibrates
 * achievable HBM bandwidth and the traffic counters. */
int c3d_copy_bandwidth(c3d_ctx *ctx, size_t nbytes, int iters, double *ms);
const char *c3d_stage_name(int i);

/* ---- Immersed boundary ----------------------------------------------
 * Marker arrays follow src/global.f90:228-247: xb, yb, zb, sb [nb];
 * normals, ub [3nb] component-major (x block, y block, z block).  Markers
 * are grouped in nzb z slabs of nxb; the first nxb1 of a slab belong to
 * body 1, the rest to body 2 (the remove_mean_per_body grouping of
 * src/immersed_boundary_geometry.f90:417-441); nxb1 = nxb is a single body
 * (center_wall, :410-411).  Every 5x5x5 stencil must
 * stay clear of the wall ghost rows. */

/* setup_IB_operators (src/immersed_boundary_operators.f90:19-256) followed
 * by compute_heaviside (src/heaviside.f90:20-54): builds the four delta
 * stencils on the device, the spreading map and the Heaviside fields and
 * coefficients.  Once per context. */
int c3d_ib_setup(c3d_ctx *ctx, int nb, const double *xb, const double *yb, const double *zb,
                 const double *normals, const double *sb, const double *ub, int nzb, int nxb,
                 int nxb1);
/* c3d_ib_setup with the per-marker y_ref_index (src/global.f90:247; the
 * y pivots of src/immersed_boundary_operators.f90:35-37 count from face
 * yref[l]); yref NULL means 1 for every marker.  Any grid type: the body's
 * stencils must stay on interior rows (y pivots measured in dymin, so a
 * body belongs in a region of uniform spacing). */
int c3d_ib_setup_ref(c3d_ctx *ctx, int nb, const double *xb, const double *yb, const double *zb,
                     const double *normals, const double *sb, const double *ub, const int *yref,
                     int nzb, int nxb, int nxb1);

/* Moving wall bodies: setup_IB_geometry for body 1 = standing_wave
 * (src/immersed_boundary_geometry.f90:176-268) or 2 = traveling_wave_x
 * (:270-385) with nxb markers along x per wall and nzb along z
 * (nb = 2 nxb nzb, the bottom wall y_ref_index 1, the top wall ny), the
 * walls at y = 0 and y = Ly of the descriptor; p1..p3 = body_param_1..3,
 * ramp = body_ramp_up_time.  The markers, the stencils, the spreading maps
 * and the Heaviside fields are all built on the device at time t
 * (setup_IB_geometry + setup_IB_operators + compute_heaviside).  Errors: the
 * standing wave's amplitude above desc.min_buffer_width (:177), a marker
 * outside [y(3), y(ny-2)] (:199-215), a stencil reaching a wall ghost row. */
int c3d_ib_setup_wave(c3d_ctx *ctx, int body, int nxb, int nzb, double p1, double p2, double p3,
                      double ramp, double t);
/* Moves the wall body to time t: the per-stage re-set-up of
 * src/time_integration.f90:110-115 (c3d_time_step_rk3 calls it at every
 * stage time t0 + rk_t(stage) dt).  Drops cached iteration graphs and the
 * history start, whose Schur operator no longer applies. */
int c3d_ib_move(c3d_ctx *ctx, double t);
/* Marker arrays as the device holds them (any pointer may be NULL). */
int c3d_ib_get_markers(c3d_ctx *ctx, double *xb, double *yb, double *zb, double *normals,
                       double *sb, double *ub, int *yref);
/* Simulation time t (src/initialization.f90:39; advanced by dt per
 * c3d_time_step_rk3, src/time_integration.f90:91,108,137,169). */
int c3d_set_time(c3d_ctx *ctx, double t);
int c3d_get_time(const c3d_ctx *ctx, double *t);

/* ---- Snapshots -------------------------------------------------------
 * The reference's big-endian stream file (output_data,
 * src/input_output.f90:498-728; tools/file_io.py:141-212): int32 length +
 * FP64 x, y, z, xm, ym, zm; int32 dims + U, V, W, P, Hu_interior,
 * Hv_interior, Hw_interior in the reference's global shapes; FP64 t, dt,
 * dpdx, nu; xb, yb, zb (int32 length + FP64); int32 nxb, nzb (per wall for
 * the wave bodies); fb, ub, sb, normals, tangents_1, tangents_2 (int32
 * length + FP64).  Without a body the masks are zeros and the body arrays
 * empty.  tangents_1/2 [3nb] come from the host's geometry (NULL: zeros;
 * the solver never reads them). */
int c3d_write_snapshot(c3d_ctx *ctx, const char *path, double dt, const double *tangents_1,
                       const double *tangents_2);
/* read_input_data (src/input_output.f90:335-489): the mesh sizes must match
 * and the coordinates equal this context's grid (to 1e-12 of each axis'
 * span); U, V, W are restored as the reference does.  what adds, beyond
 * the reference (whose restart loses them): C3D_RESTORE_P the pressure,
 * C3D_RESTORE_STATE t and dPdx (a moving body is moved to t),
 * C3D_RESTORE_FORCE fb, the warm start of the next IB solve (the body set
 * up first, same nb; nothing to do for a body-free file and context).
 * With all three a restart continues bit for bit.
 * dt_out (may be NULL) receives the recorded dt. */
#define C3D_RESTORE_P 1
#define C3D_RESTORE_STATE 2
#define C3D_RESTORE_FORCE 4
#define C3D_RESTORE_ALL 7
int c3d_read_snapshot(c3d_ctx *ctx, const char *path, int what, double *dt_out);

/* Stencil s (0 u, 1 v, 2 w, 3 centre) as the reference stores it
 * (src/immersed_boundary_operators.f90:46-244): for marker l and point
 * p = i + 5*(j + 5*k), entry [l*125 + p] of the Fortran node indices, the
 * weight dx*dymin*dz*phi_x*phi_y*phi_z and dxn = n . (x_node - xb).
 * Arrays of nb*125; any may be NULL. */
int c3d_ib_get_stencil(c3d_ctx *ctx, int s, int *xi, int *yi, int *zi, double *w, double *dxn);
/* Heaviside outputs: E1nHc_exterior [nb], E1nH_exterior [3nb], the
 * interior masks Hu/Hv/Hw_interior (field shapes) and Hc_exterior (rhs_p
 * shape).  Any may be NULL. */
int c3d_ib_get_heaviside(c3d_ctx *ctx, double *E1nHc, double *E1nH, double *Hu_int,
                         double *Hv_int, double *Hw_int, double *Hc_ext);
/* regT (with_dxn 0) / regT_1n (1) of the resident U, V, W -> out[3nb]
 * (src/immersed_boundary_operators.f90:296-342). */
int c3d_ib_interpolate(c3d_ctx *ctx, int with_dxn, double *out);
/* regT (with_dxn 0) / regT_1n (1) of host arrays U_, V_, W_ -> out[3nb]
 * (src/immersed_boundary_operators.f90:296-342, the Eu_ argument); the
 * resident velocity is not touched. */
int c3d_ib_interpolate_arrays(c3d_ctx *ctx, int with_dxn, const double *U, const double *V,
                              const double *W, double *out);
/* compute_heaviside (src/heaviside.f90:20-54) again for the current markers:
 * Hu/Hv/Hw_interior, Hc_exterior, E1nHc_exterior and E1nH_exterior
 * (c3d_ib_setup* and c3d_ib_move already run it). */
int c3d_ib_compute_heaviside(c3d_ctx *ctx);
/* regTc_1n of a host centre field p (rhs_p shape) -> out[nb] (:344-357). */
int c3d_ib_interpolate_c(c3d_ctx *ctx, const double *p, double *out);
/* regu / regv / regw (comp 0, 1, 2) of f[nb] into the host field F of that
 * component's shape, zero outside the support (:359-400, :544-603). */
int c3d_ib_spread(c3d_ctx *ctx, int comp, const double *f, double *F);
/* remove_mean_per_body of f[nb] in place
 * (src/immersed_boundary_geometry.f90:400-442). */
int c3d_ib_remove_mean(c3d_ctx *ctx, double *f);
/* schur (src/projection.f90:98-144): Sf = S f for host f, Sf [4nb].  The
 * resident velocity is not used as scratch (the reference overwrites it). */
int c3d_ib_schur(c3d_ctx *ctx, const double *f, double *Sf, double dt, int rk_step);
/* bicgstab (src/projection.f90:146-186) on host arrays: x [4nb] is the
 * initial guess and the solution.  *iters, *err (= ||r||^2) may be NULL.
 * Returns C3D_ENOCONV if maxit was reached above tolerance (x is still
 * written back). */
int c3d_ib_bicgstab(c3d_ctx *ctx, double *x, const double *b, double tol, int maxit, double dt,
                    int rk_step, int *iters, double *err);
/* Solver settings for the IB projection: absolute tolerance cg_tol,
 * cg_max_iter, and the initial guess: guess_mode 0 = the previous force
 * fb (the reference's warm start), 1 = zero, 2 = the force last set by
 * c3d_ib_set_force for the next solve only, then 0, 3 = the previous force
 * plus the learned increment (c3d_mlp_*), 4 = the least-squares
 * combination of the last 8 converged forces of the same dt, fitted through
 * their right-hand sides (for a static body the Schur operator depends on dt
 * only; a moving body clears the stored forces at every move; the previous
 * force while none is stored), 5 = the converged force of the same RK3
 * stage (rk_step) of the previous step (the previous force until one is
 * stored).  Modes 3-5 are additions with no counterpart in the reference;
 * BiCGSTAB starts from the true residual in every mode. */
int c3d_ib_set_solver(c3d_ctx *ctx, double tol, int maxit, int guess_mode);
/* on != 0 (the default): each BiCGSTAB iteration of the IB solves is
 * replayed from a HIP graph captured once per (solution vector, dt, rk_step
 * == 1); 0: launched kernel by kernel.  Both run the same kernels in the same
 * order and give identical results.  The default at c3d_ib_setup is on
 * unless the environment sets C3D_IB_GRAPHS=0; c3d_ib_setup_wave starts
 * with it off (each move would force a recapture).  No reference counterpart
 * (launch mechanics of src/projection.f90:160-182). */
int c3d_ib_set_graphs(c3d_ctx *ctx, int on);
/* Resident IB force fb = (dudn_jump[3nb], p_jump[nb]) [4nb]. */
int c3d_ib_set_force(c3d_ctx *ctx, const double *fb);
int c3d_ib_get_force(c3d_ctx *ctx, double *fb);
/* compute_IB_projection (src/projection.f90:48-96) on the resident state,
 * after c3d_compute_non_ib_projection (which keeps U_interim when a body is
 * set up) and boundary conditions.  Returns C3D_ENOCONV with the fields
 * updated if BiCGSTAB stopped at its iteration limit. */
int c3d_compute_ib_projection(c3d_ctx *ctx, double dt, int rk_step, int *iters, double *err);
/* check_slip (src/monitor.f90:223-231): max|E u - E1nH_exterior*dudn - ub|. */
int c3d_check_slip(c3d_ctx *ctx, double *max_slip);
/* BiCGSTAB statistics: iterations of the last solve, iterations summed over
 * all solves (cg_accum_iter, src/projection.f90:181), last ||r||^2. */
int c3d_ib_stats(c3d_ctx *ctx, int *last_iters, long long *accum_iters, double *last_err);
/* exterior_pressure_gradient (src/input_output.f90:52, src/equations.f90:147-151):
 * 1 multiplies the driving pressure gradient by H_interior, 0 applies it
 * everywhere.  The bulk-flow integral is H_interior-masked once a body is
 * set up (src/mass_flow.f90:24-53). */
int c3d_ib_set_exterior_pressure_gradient(c3d_ctx *ctx, int flag);

/* ---- Learned initial guess for the Schur solve --------------------------
 * No reference counterpart: the reference starts BiCGSTAB from the previous
 * force (src/projection.f90:60).  A 54 -> 256 -> 256 -> 256 -> 4 ReLU network
 * (BF16 matrix-core GEMMs, FP32 accumulation and master weights) maps, for
 * each marker, 6 channels over its 3x3 (azimuth x span) neighbourhood -- the
 * residual of the warm start r = b - S fprev (velocity rows x, y, z and the
 * pressure row) and the previous force (dudn_x, dudn_y) -- to a force
 * increment d (dudn_x, dudn_y, dudn_z, p_jump).  Inputs are divided by
 * iscale[6], outputs multiplied by oscale[4].  c3d_ib_set_solver guess_mode
 * 3 starts BiCGSTAB from fb = fprev + a d with the step a minimising
 * |b - S(fprev + a d)|, then iterates to the same tolerance. */
int c3d_mlp_param_count(void);
/* (Re)initialises the weights (He-uniform, SplitMix64 seed), zero biases,
 * unit scales, and clears the Adam state. */
int c3d_mlp_init(c3d_ctx *ctx, unsigned long long seed);
int c3d_mlp_set_scales(c3d_ctx *ctx, const float *iscale, const float *oscale);
/* Flat FP32 parameters: per layer W[out][in] then b[out]. */
int c3d_mlp_get_params(c3d_ctx *ctx, float *params);
int c3d_mlp_set_params(c3d_ctx *ctx, const float *params);
/* Y[n][4] = network(X[n][54]) (scaled units). */
int c3d_mlp_forward(c3d_ctx *ctx, const float *X, int n, float *Y);
/* Mean-squared-error gradient of the parameters on (X, T[n][4]) without an
 * update; *loss (may be NULL) receives the loss. */
int c3d_mlp_gradients(c3d_ctx *ctx, const float *X, const float *T, int n, float *grad,
                      float *loss);
/* Adam training on the GPU: `epochs` passes over the n samples, shuffled
 * per epoch (SplitMix64 seed), minibatches of `batch` (multiple of 16);
 * loss_hist[epochs] (may be NULL) receives the full-set loss per epoch. */
int c3d_mlp_train(c3d_ctx *ctx, const float *X, const float *T, int n, int epochs, int batch,
                  float lr, unsigned long long seed, float *loss_hist);
/* C[M][N] = op(A) op(B) with the network's BF16 matrix-core GEMM; A is
 * M x K (K x M if ta), B is K x N (N x K if tb), row-major.  Test hook. */
int c3d_mlp_gemm_test(c3d_ctx *ctx, int M, int N, int K, int ta, int tb, const float *A,
                      const float *B, float *C);
/* Training pair of the last IB solve: X[nb][54] unscaled features (the
 * residual of the previous force and that force, as the network sees them
 * before scaling) and dF[4nb] = converged force minus the previous force
 * (component-major).  Costs one Schur product. */
int c3d_ib_last_pair(c3d_ctx *ctx, float *X, double *dF);

#ifdef __cplusplus
}
#endif
#endif /* C3D_H */
