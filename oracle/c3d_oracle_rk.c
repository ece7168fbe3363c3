/*
 * c3d_oracle_rk.c -- CPU FP64 restatement of the RK3 advection-diffusion
 * right-hand sides, the RK3 stage update, compute_dt and the constant
 * x-mass-flow correction.
 *
 * TEST INFRASTRUCTURE ONLY (see c3d_oracle.h).
 *
 * The reference computes each right-hand side as a sum of whole-array
 * passes through the scratch arrays term_1/term_2 (src/equations.f90).
 * Here each term is evaluated pointwise with the same operands, the same
 * operation order inside each term, and the same order of accumulation
 * into rhs (convective x, y, z, then viscous, then the pressure gradient),
 * so the result equals the array-pass formulation exactly.  x**2d0 is
 * evaluated as x*x.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "c3d_oracle.h"
#include "c3d_oracle_impl.h"

/* grid access through the shared internal struct */
typedef struct orc_grid G;
static void gload(const orc_grid *g, G *q) { *q = *g; }
static void gfree(G *q) { (void)q; }

/* 1-based Fortran accessors */
#define X(i) q.x[(i) - 1]
#define Y(j) q.y[(j) - 1]
#define Z(k) q.z[(k) - 1]
#define XG(i) q.xg[(i) - 1]
#define YG(j) q.yg[(j) - 1]
#define ZG(k) q.zg[(k) - 1]
#define IX(i, j, k, n1, n2) ((size_t)((i) - 1) + (size_t)(n1) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))
#define Uf(i, j, k) U[IX(i, j, k, q.nx, q.nyg)]
#define Vf(i, j, k) V[IX(i, j, k, q.nxg, q.ny)]
#define Wf(i, j, k) W[IX(i, j, k, q.nxg, q.nyg)]

/* compute_rhs_u: src/equations.f90:24-153.  F has bounds
 * (2:nx-1, 2:nyg-1, 2:nzg-1).  Hu may be NULL (H = 1). */
void orc_rhs_u(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hu, double nu, double dPdx, double *F) {
  G q;
  gload(g, &q);
  const int n1 = q.nx - 2, n2 = q.nyg - 2;
  for (int k = 2; k <= q.nzg - 1; k++)
    for (int j = 2; j <= q.nyg - 1; j++)
      for (int i = 2; i <= q.nx - 1; i++) {
        /* 1) -d(u^2)/dx with u averaged to centres (:43-57) */
        double uc1 = 0.5 * (Uf(i - 1, j, k) + Uf(i, j, k));
        double uc2 = 0.5 * (Uf(i, j, k) + Uf(i + 1, j, k));
        double a = -(uc2 * uc2 - uc1 * uc1) / (XG(i + 1) - XG(i));
        /* 2) -d(uv)/dy at y faces j and j-1 (:62-79) */
        double uv1 = (0.5 * (Uf(i, j, k) + Uf(i, j + 1, k))) * (0.5 * (Vf(i, j, k) + Vf(i + 1, j, k)));
        double uv0 = (0.5 * (Uf(i, j - 1, k) + Uf(i, j, k))) * (0.5 * (Vf(i, j - 1, k) + Vf(i + 1, j - 1, k)));
        double b = -(uv1 - uv0) / (Y(j) - Y(j - 1));
        /* 3) -d(uw)/dz at z faces k and k-1 (:84-101) */
        double uw1 = (0.5 * (Uf(i, j, k) + Uf(i, j, k + 1))) * (0.5 * (Wf(i, j, k) + Wf(i + 1, j, k)));
        double uw0 = (0.5 * (Uf(i, j, k - 1) + Uf(i, j, k))) * (0.5 * (Wf(i, j, k - 1) + Wf(i + 1, j, k - 1)));
        double c = -(uw1 - uw0) / (Z(k) - Z(k - 1));
        /* 4) viscous term (:108-144) */
        double dx_1 = X(i) - X(i - 1), dx_2 = X(i + 1) - X(i), dx_3 = XG(i + 1) - XG(i);
        double dy_3 = Y(j) - Y(j - 1);
        double dz_1 = ZG(k) - ZG(k - 1), dz_2 = ZG(k + 1) - ZG(k), dz_3 = Z(k) - Z(k - 1);
        double t1 = (Uf(i, j + 1, k) - Uf(i, j, k)) / (YG(j + 1) - YG(j));
        double t1m = (Uf(i, j, k) - Uf(i, j - 1, k)) / (YG(j) - YG(j - 1));
        double visc = 1.0 / dx_3 * (nu * 1.0 / dx_2 * (Uf(i + 1, j, k) - Uf(i, j, k)) -
                                    nu * 1.0 / dx_1 * (Uf(i, j, k) - Uf(i - 1, j, k))) +
                      1.0 / dy_3 * (nu * t1 - nu * t1m) +
                      1.0 / dz_3 * (nu * 1.0 / dz_2 * (Uf(i, j, k + 1) - Uf(i, j, k)) -
                                    nu * 1.0 / dz_1 * (Uf(i, j, k) - Uf(i, j, k - 1)));
        double h = Hu ? Hu[IX(i, j, k, q.nx, q.nyg)] : 1.0;
        double r = a;
        r = r + b;
        r = r + c;
        r = r + visc;
        r = r + dPdx * h; /* :147-151, exterior_pressure_gradient */
        F[(size_t)(i - 2) + (size_t)n1 * ((size_t)(j - 2) + (size_t)n2 * (size_t)(k - 2))] = r;
      }
  gfree(&q);
}

/* compute_rhs_v: src/equations.f90:160-288.  F bounds (2:nxg-1, 2:ny-1, 2:nzg-1). */
void orc_rhs_v(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hv, double nu, double dPdy, double *F) {
  G q;
  gload(g, &q);
  const int n1 = q.nxg - 2, n2 = q.ny - 2;
  for (int k = 2; k <= q.nzg - 1; k++)
    for (int j = 2; j <= q.ny - 1; j++)
      for (int i = 2; i <= q.nxg - 1; i++) {
        /* 1) -d(v^2)/dy with v averaged to centres (:179-192) */
        double vc1 = 0.5 * (Vf(i, j - 1, k) + Vf(i, j, k));
        double vc2 = 0.5 * (Vf(i, j, k) + Vf(i, j + 1, k));
        double a = -(vc2 * vc2 - vc1 * vc1) / (YG(j + 1) - YG(j));
        /* 2) -d(uv)/dx at x faces i and i-1 (:197-214) */
        double uv1 = (0.5 * (Uf(i, j, k) + Uf(i, j + 1, k))) * (0.5 * (Vf(i, j, k) + Vf(i + 1, j, k)));
        double uv0 = (0.5 * (Uf(i - 1, j, k) + Uf(i - 1, j + 1, k))) * (0.5 * (Vf(i - 1, j, k) + Vf(i, j, k)));
        double b = -(uv1 - uv0) / (X(i) - X(i - 1));
        /* 3) -d(vw)/dz at z faces k and k-1 (:219-236) */
        double vw1 = (0.5 * (Vf(i, j, k) + Vf(i, j, k + 1))) * (0.5 * (Wf(i, j, k) + Wf(i, j + 1, k)));
        double vw0 = (0.5 * (Vf(i, j, k - 1) + Vf(i, j, k))) * (0.5 * (Wf(i, j, k - 1) + Wf(i, j + 1, k - 1)));
        double c = -(vw1 - vw0) / (Z(k) - Z(k - 1));
        /* 4) viscous term (:245-278) */
        double dx_1 = XG(i) - XG(i - 1), dx_2 = XG(i + 1) - XG(i), dx_3 = X(i) - X(i - 1);
        double dy_1 = Y(j) - Y(j - 1), dy_2 = Y(j + 1) - Y(j), dy_3 = YG(j + 1) - YG(j);
        double dz_1 = ZG(k) - ZG(k - 1), dz_2 = ZG(k + 1) - ZG(k), dz_3 = Z(k) - Z(k - 1);
        double visc = 1.0 / dx_3 * (nu * 1.0 / dx_2 * (Vf(i + 1, j, k) - Vf(i, j, k)) -
                                    nu * 1.0 / dx_1 * (Vf(i, j, k) - Vf(i - 1, j, k))) +
                      1.0 / dy_3 * (nu * 1.0 / dy_2 * (Vf(i, j + 1, k) - Vf(i, j, k)) -
                                    nu * 1.0 / dy_1 * (Vf(i, j, k) - Vf(i, j - 1, k))) +
                      1.0 / dz_3 * (nu * 1.0 / dz_2 * (Vf(i, j, k + 1) - Vf(i, j, k)) -
                                    nu * 1.0 / dz_1 * (Vf(i, j, k) - Vf(i, j, k - 1)));
        double h = Hv ? Hv[IX(i, j, k, q.nxg, q.ny)] : 1.0;
        double r = a;
        r = r + b;
        r = r + c;
        r = r + visc;
        r = r + dPdy * h;
        F[(size_t)(i - 2) + (size_t)n1 * ((size_t)(j - 2) + (size_t)n2 * (size_t)(k - 2))] = r;
      }
  gfree(&q);
}

/* compute_rhs_w: src/equations.f90:295-423.  F bounds (2:nxg-1, 2:nyg-1, 2:nz-1). */
void orc_rhs_w(const orc_grid *g, const double *U, const double *V, const double *W,
               const double *Hw, double nu, double dPdz, double *F) {
  G q;
  gload(g, &q);
  const int n1 = q.nxg - 2, n2 = q.nyg - 2;
  for (int k = 2; k <= q.nz - 1; k++)
    for (int j = 2; j <= q.nyg - 1; j++)
      for (int i = 2; i <= q.nxg - 1; i++) {
        /* 1) -d(w^2)/dz with w averaged to centres (:314-327) */
        double wc1 = 0.5 * (Wf(i, j, k - 1) + Wf(i, j, k));
        double wc2 = 0.5 * (Wf(i, j, k) + Wf(i, j, k + 1));
        double a = -(wc2 * wc2 - wc1 * wc1) / (ZG(k + 1) - ZG(k));
        /* 2) -d(uw)/dx at x faces i and i-1 (:332-349) */
        double uw1 = (0.5 * (Uf(i, j, k) + Uf(i, j, k + 1))) * (0.5 * (Wf(i, j, k) + Wf(i + 1, j, k)));
        double uw0 = (0.5 * (Uf(i - 1, j, k) + Uf(i - 1, j, k + 1))) * (0.5 * (Wf(i - 1, j, k) + Wf(i, j, k)));
        double b = -(uw1 - uw0) / (X(i) - X(i - 1));
        /* 3) -d(vw)/dy at y faces j and j-1 (:354-371) */
        double vw1 = (0.5 * (Vf(i, j, k) + Vf(i, j, k + 1))) * (0.5 * (Wf(i, j, k) + Wf(i, j + 1, k)));
        double vw0 = (0.5 * (Vf(i, j - 1, k) + Vf(i, j - 1, k + 1))) * (0.5 * (Wf(i, j - 1, k) + Wf(i, j, k)));
        double c = -(vw1 - vw0) / (Y(j) - Y(j - 1));
        /* 4) viscous term (:376-413) */
        double dx_1 = XG(i) - XG(i - 1), dx_2 = XG(i + 1) - XG(i), dx_3 = X(i) - X(i - 1);
        double dy_3 = Y(j) - Y(j - 1);
        double dz_1 = Z(k) - Z(k - 1), dz_2 = Z(k + 1) - Z(k), dz_3 = ZG(k + 1) - ZG(k);
        double t1 = (Wf(i, j + 1, k) - Wf(i, j, k)) / (YG(j + 1) - YG(j));
        double t1m = (Wf(i, j, k) - Wf(i, j - 1, k)) / (YG(j) - YG(j - 1));
        double visc = 1.0 / dx_3 * (nu * 1.0 / dx_2 * (Wf(i + 1, j, k) - Wf(i, j, k)) -
                                    nu * 1.0 / dx_1 * (Wf(i, j, k) - Wf(i - 1, j, k))) +
                      1.0 / dy_3 * (nu * t1 - nu * t1m) +
                      1.0 / dz_3 * (nu * 1.0 / dz_2 * (Wf(i, j, k + 1) - Wf(i, j, k)) -
                                    nu * 1.0 / dz_1 * (Wf(i, j, k) - Wf(i, j, k - 1)));
        double h = Hw ? Hw[IX(i, j, k, q.nxg, q.nyg)] : 1.0;
        double r = a;
        r = r + b;
        r = r + c;
        r = r + visc;
        r = r + dPdz * h;
        F[(size_t)(i - 2) + (size_t)n1 * ((size_t)(j - 2) + (size_t)n2 * (size_t)(k - 2))] = r;
      }
  gfree(&q);
}

/* RK3 stage update, src/time_integration.f90:105-107,134-136,163-168:
 *   U(int) = Uo(int) + dt*rk_coef(1,1)*F1                      (stage 1)
 *   U(int) = Uo(int) + dt*( c21*F1 + c22*F2 )                  (stage 2)
 *   U(int) = Uo(int) + dt*( c31*F1 + c32*F2 + c33*F3 )         (stage 3)
 * with the Wray coefficients of src/initialization.f90:406-418.
 * which: 0 U, 1 V, 2 W (selects the interior ranges). */
void orc_rk_update(const orc_grid *g, int which, int stage, double dt, const double *Fo1,
                   const double *Fo2, const double *Fo3, const double *Xo, double *Xn) {
  G q;
  gload(g, &q);
  int n1, n2, i1, j1, k1;
  if (which == 0) { n1 = q.nx; n2 = q.nyg; i1 = q.nx - 1; j1 = q.nyg - 1; k1 = q.nzg - 1; }
  else if (which == 1) { n1 = q.nxg; n2 = q.ny; i1 = q.nxg - 1; j1 = q.ny - 1; k1 = q.nzg - 1; }
  else { n1 = q.nxg; n2 = q.nyg; i1 = q.nxg - 1; j1 = q.nyg - 1; k1 = q.nz - 1; }
  const double c11 = 8.0 / 15.0, c21 = 1.0 / 4.0, c22 = 5.0 / 12.0, c31 = 1.0 / 4.0, c32 = 0.0,
               c33 = 3.0 / 4.0;
  const int m1 = i1 - 1, m2 = j1 - 1;
  for (int k = 2; k <= k1; k++)
    for (int j = 2; j <= j1; j++)
      for (int i = 2; i <= i1; i++) {
        size_t f = (size_t)(i - 2) + (size_t)m1 * ((size_t)(j - 2) + (size_t)m2 * (size_t)(k - 2));
        size_t o = IX(i, j, k, n1, n2);
        double v;
        if (stage == 1) v = Xo[o] + dt * c11 * Fo1[f];
        else if (stage == 2) v = Xo[o] + dt * (c21 * Fo1[f] + c22 * Fo2[f]);
        else v = Xo[o] + dt * (c31 * Fo1[f] + c32 * Fo2[f] + c33 * Fo3[f]);
        Xn[o] = v;
      }
  gfree(&q);
}

/* compute_dt: src/time_integration.f90:219-267 (maximum of dx/|u|, as written). */
double orc_compute_dt(const orc_grid *g, const double *U, const double *V, const double *W,
                      double cfl, double nu) {
  G q;
  gload(g, &q);
  double lU = 0.0, lV = 0.0, lW = 0.0;
  for (int i = 2; i <= q.nxg - 1; i++)
    for (int j = 2; j <= q.nyg - 1; j++)
      for (int k = 2; k <= q.nzg - 1; k++) {
        lU = fmax(lU, (XG(i + 1) - XG(i)) / fabs(Uf(i, j, k)));
        lV = fmax(lV, (YG(j + 1) - YG(j)) / fabs(Vf(i, j, k)));
        lW = fmax(lW, (ZG(k + 1) - ZG(k)) / fabs(Wf(i, j, k)));
      }
  double cu = cfl * lU, cv = cfl * lV, cw = cfl * lW;
  double dtc = fmin(fmin(cu, cv), cw);
  double vu = cfl * (q.dxmin * q.dxmin) / nu, vv = cfl * (q.dymin * q.dymin) / nu,
         vw = cfl * (q.dzmin * q.dzmin) / nu;
  double dtv = fmin(fmin(vu, vv), vw);
  double dt = fmin(dtc, dtv);
  dt = fmin(dt, 0.1);
  if (cfl < 0) dt = -cfl;
  gfree(&q);
  return dt;
}

/* compute_mean_mass_flow_U / compute_dPx_for_constant_mass_flow integral:
 * src/mass_flow.f90:24-53,144-175.  Returns Int_U/norm (Hu may be NULL). */
double orc_mean_mass_flow_u(const orc_grid *g, const double *U, const double *Hu) {
  G q;
  gload(g, &q);
  double s = 0.0, nrm = 0.0;
  for (int k = 2; k <= q.nzg - 1; k++)
    for (int j = 3; j <= q.nyg - 1; j++)
      for (int i = 2; i <= q.nx - 1; i++) {
        double h1 = Hu ? Hu[IX(i, j, k, q.nx, q.nyg)] : 1.0;
        double h0 = Hu ? Hu[IX(i, j - 1, k, q.nx, q.nyg)] : 1.0;
        s = s + (h1 * Uf(i, j, k) + h0 * Uf(i, j - 1, k)) * 0.5 * (YG(j) - YG(j - 1));
        nrm = nrm + (h1 + h0) * 0.5 * (YG(j) - YG(j - 1));
      }
  gfree(&q);
  return s / nrm;
}

/* output_monitor's field summaries (src/monitor.f90:27-39, 57-59) on one
 * domain: out[0..2] = Maxval of U(2:nx-1,2:nyg-1,2:nzg-1),
 * V(2:nxg-1,2:ny-1,2:nzg-1), W(2:nxg-1,2:nyg-1,2:nz-1); out[3..5] = their
 * Sums divided by nxm*nym*nzm (nxm = nx-1 etc., src/initialization.f90:79-81);
 * out[6] = count of non-finite entries in the three interiors. */
void orc_monitor(const orc_grid *g, const double *U, const double *V, const double *W,
                 double *out) {
  G q;
  gload(g, &q);
  double s[3] = {0, 0, 0}, m[3] = {-INFINITY, -INFINITY, -INFINITY}, bad = 0;
  for (int k = 2; k <= q.nzg - 1; k++)
    for (int j = 2; j <= q.nyg - 1; j++)
      for (int i = 2; i <= q.nx - 1; i++) {
        const double v = Uf(i, j, k);
        s[0] += v; m[0] = fmax(m[0], v); bad += isfinite(v) ? 0 : 1;
      }
  for (int k = 2; k <= q.nzg - 1; k++)
    for (int j = 2; j <= q.ny - 1; j++)
      for (int i = 2; i <= q.nxg - 1; i++) {
        const double v = Vf(i, j, k);
        s[1] += v; m[1] = fmax(m[1], v); bad += isfinite(v) ? 0 : 1;
      }
  for (int k = 2; k <= q.nz - 1; k++)
    for (int j = 2; j <= q.nyg - 1; j++)
      for (int i = 2; i <= q.nxg - 1; i++) {
        const double v = Wf(i, j, k);
        s[2] += v; m[2] = fmax(m[2], v); bad += isfinite(v) ? 0 : 1;
      }
  const double vol = (double)(q.nx - 1) * (double)(q.ny - 1) * (double)(q.nz - 1);
  for (int c = 0; c < 3; c++) {
    out[c] = m[c];
    out[3 + c] = s[c] / vol;
  }
  out[6] = bad;
  gfree(&q);
}

/* Plane means of compute_statistics on one domain (src/statistics.f90:41-111,
 * last-rank branch :64-73): term_1 = interpolate_x(U) (0.5*(U(i-1) + U(i))
 * at centre i, src/interpolation.f90:21-36), term_2 = interpolate_y(V)
 * (0.5*(V(j-1) + V(j)) at centre j, :44-58) with term_2(:,1,:) =
 * -term_2(:,2,:) and term_2(:,nyg,:) = -term_2(:,nyg-1,:) (:47-49); sums
 * over i = 2..nxg-2, k = 2..nzg-2 (W: k = 2..nz-1) in Fortran's array order,
 * divided as in :103-111.  out[q*nyg + j-1], q = 0 Umean, 1 Vmean, 2 Wmean,
 * 3 U2mean, 4 V2mean, 5 W2mean, 6 UVmean. */
void orc_plane_statistics(const orc_grid *g, const double *U, const double *V, const double *W,
                          double *out) {
  G q;
  gload(g, &q);
  const int n = q.nyg;
  const double nuv = (double)((q.nxg - 3) * (q.nzg - 3)), nw = (double)((q.nxg - 3) * (q.nz - 2));
  for (int j = 1; j <= n; j++) {
    const int jv = (j == 1) ? 2 : (j == n ? n - 1 : j);
    const double sg = (j == 1 || j == n) ? -1.0 : 1.0;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 2; k <= q.nzg - 2; k++)
      for (int i = 2; i <= q.nxg - 2; i++) {
        const double t1 = 0.5 * (Uf(i - 1, j, k) + Uf(i, j, k));
        const double t2 = sg * (0.5 * (Vf(i, jv - 1, k) + Vf(i, jv, k)));
        s[0] += t1;
        s[1] += t2;
        s[3] += t1 * t1;
        s[4] += t2 * t2;
        s[6] += t1 * t2;
      }
    for (int k = 2; k <= q.nz - 1; k++)
      for (int i = 2; i <= q.nxg - 2; i++) {
        s[2] += Wf(i, j, k);
        s[5] += Wf(i, j, k) * Wf(i, j, k);
      }
    for (int c = 0; c < 7; c++) out[(size_t)c * n + j - 1] = s[c] / ((c == 2 || c == 5) ? nw : nuv);
  }
  gfree(&q);
}

/* Wall shear of compute_statistics on one domain (src/statistics.f90:41-131):
 * plane means over the unique cells of the x-centred U (interpolate_x,
 * src/interpolation.f90:21-36: 0.5*(U(i-1) + U(i)) at centre i) over
 * i = 2..nxg-2, k = 2..nzg-2, and of W over i = 2..nxg-2, k = 2..nz-1
 * (the last-rank branch, :64-73), divided by (nxg-3)*(nzg-3) resp.
 * (nxg-3)*(nz-2) (:103-105); wall gradients from rows 1, 2 and nyg-1, nyg
 * (:114-117); the wall Reynolds stresses are zero (:120-123);
 * utau = ((UV_t - UV_b - nu*dUdy_t + nu*dUdy_b)/Ly)**0.5 (:126-127).
 * Sums run in Fortran's array order (i fastest).  Umean, Wmean [nyg] may
 * be NULL. */
void orc_wall_shear(const orc_grid *g, const double *U, const double *W, double nu,
                    double *utau, double *wtau, double *Umean, double *Wmean) {
  G q;
  gload(g, &q);
  double *um = malloc((size_t)q.nyg * sizeof(double));
  double *wm = malloc((size_t)q.nyg * sizeof(double));
  for (int j = 1; j <= q.nyg; j++) {
    double su = 0.0, sw = 0.0;
    for (int k = 2; k <= q.nzg - 2; k++)
      for (int i = 2; i <= q.nxg - 2; i++) su += 0.5 * (Uf(i - 1, j, k) + Uf(i, j, k));
    for (int k = 2; k <= q.nz - 1; k++)
      for (int i = 2; i <= q.nxg - 2; i++) sw += Wf(i, j, k);
    um[j - 1] = su / (double)((q.nxg - 3) * (q.nzg - 3));
    wm[j - 1] = sw / (double)((q.nxg - 3) * (q.nz - 2));
  }
  const int n = q.nyg;
  const double dub = (um[1] - um[0]) / (YG(2) - YG(1));
  const double dut = (um[n - 1] - um[n - 2]) / (YG(n) - YG(n - 1));
  const double dwb = (wm[1] - wm[0]) / (YG(2) - YG(1));
  const double dwt = (wm[n - 1] - wm[n - 2]) / (YG(n) - YG(n - 1));
  const double uvt = 0.0, uvb = 0.0, vwt = 0.0, vwb = 0.0;
  *utau = sqrt((uvt - uvb - nu * dut + nu * dub) / q.Ly);
  *wtau = sqrt((vwt - vwb - nu * dwt + nu * dwb) / q.Ly);
  if (Umean) memcpy(Umean, um, (size_t)n * sizeof(double));
  if (Wmean) memcpy(Wmean, wm, (size_t)n * sizeof(double));
  free(um);
  free(wm);
  gfree(&q);
}

/* compute_time_step_RK3 without a body: src/time_integration.f90:86-193.
 * Per stage: right-hand sides, stage update, boundary conditions,
 * projection, boundary conditions; after stage 3 the constant x-mass-flow
 * correction when q0 is finite (x_mass_cte == 1):
 *   dPdx_prev = dPdx; dPdx = Q0 - <U>; U(int) += dPdx;
 *   dPdx = (dPdx_prev + dPdx/dt)/2; boundary conditions.
 * *dPdx is read and updated.  P receives the pressure side effect of the
 * first-stage Poisson solve.  Pint (rhs_p shape) receives P_interim. */
void orc_time_step_rk3(const orc_grid *g, double *U, double *V, double *W, double *P,
                       double *Pint, double dt, double nu, double *dPdx, double q0,
                       int nthreads) {
  G q;
  gload(g, &q);
  size_t nu_ = (size_t)q.nx * q.nyg * q.nzg, nv = (size_t)q.nxg * q.ny * q.nzg,
         nw = (size_t)q.nxg * q.nyg * q.nz;
  size_t fu = (size_t)(q.nx - 2) * (q.nyg - 2) * (q.nzg - 2);
  size_t fv = (size_t)(q.nxg - 2) * (q.ny - 2) * (q.nzg - 2);
  size_t fw = (size_t)(q.nxg - 2) * (q.nyg - 2) * (q.nz - 2);
  size_t nr = (size_t)(q.nxg - 2) * (q.nyg - 2) * (q.nzg - 1);
  double *Uo = malloc(nu_ * 8), *Vo = malloc(nv * 8), *Wo = malloc(nw * 8);
  double *Fu[3], *Fv[3], *Fw[3];
  for (int s = 0; s < 3; s++) {
    Fu[s] = calloc(fu, 8); Fv[s] = calloc(fv, 8); Fw[s] = calloc(fw, 8);
  }
  double *rhs = malloc(nr * 8);
  memcpy(Uo, U, nu_ * 8); memcpy(Vo, V, nv * 8); memcpy(Wo, W, nw * 8);
  const double dPdy = 0.0, dPdz = 0.0;
  for (int st = 1; st <= 3; st++) {
    orc_rhs_u(g, U, V, W, NULL, nu, *dPdx, Fu[st - 1]);
    orc_rhs_v(g, U, V, W, NULL, nu, dPdy, Fv[st - 1]);
    orc_rhs_w(g, U, V, W, NULL, nu, dPdz, Fw[st - 1]);
    orc_rk_update(g, 0, st, dt, Fu[0], Fu[1], Fu[2], Uo, U);
    orc_rk_update(g, 1, st, dt, Fv[0], Fv[1], Fv[2], Vo, V);
    orc_rk_update(g, 2, st, dt, Fw[0], Fw[1], Fw[2], Wo, W);
    orc_apply_bc(g, U, V, W);
    orc_projection(g, U, V, W, NULL, NULL, NULL, Pint, rhs, P, st, dt, 0, nthreads);
    orc_apply_bc(g, U, V, W);
  }
  if (isfinite(q0)) {
    double prev = *dPdx;
    double d = q0 - orc_mean_mass_flow_u(g, U, NULL);
    for (int k = 2; k <= q.nzg - 1; k++)
      for (int j = 2; j <= q.nyg - 1; j++)
        for (int i = 2; i <= q.nx - 1; i++) Uf(i, j, k) = Uf(i, j, k) + d;
    *dPdx = 0.5 * (prev + d / dt);
    orc_apply_bc(g, U, V, W);
  }
  for (int s = 0; s < 3; s++) { free(Fu[s]); free(Fv[s]); free(Fw[s]); }
  free(Uo); free(Vo); free(Wo); free(rhs);
  gfree(&q);
}
