/*
 * c3d_oracle_ib.c -- CPU FP64 restatement of the immersed-boundary path:
 * delta kernel, operator set-up, interpolation E / E1n, spreading R,
 * per-body mean removal, Heaviside fields, Schur complement, BiCGSTAB and
 * the IB projection.
 *
 * TEST INFRASTRUCTURE ONLY (see c3d_oracle.h).
 *
 * Single-domain semantics of the reference's MPI z-slab code:
 *  - stencil indices are the reference's with the rank mapping removed; a
 *    centre z index that the reference sends to the duplicate plane nzg-1
 *    (src/immersed_boundary_operators.f90:66-71,116-121,207-212) is stored
 *    as its canonical plane 2, so spreading accumulates there (the
 *    reference's support fold-back, src/boundary_conditions.f90:737-906);
 *    interpolation reads the same value either way after the periodic
 *    boundary conditions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "c3d_oracle.h"
#include "c3d_oracle_impl.h"

#define NW 125 /* (2*2+1)^3 stencil points, src/global.f90:266-269 */
#define IX(i, j, k, n1, n2) ((size_t)((i) - 1) + (size_t)(n1) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))
#define RX(i, j, k, g) ((size_t)((i) - 2) + (size_t)((g)->nxg - 2) * ((size_t)((j) - 2) + (size_t)((g)->nyg - 2) * (size_t)((k) - 2)))

struct orc_ib {
  const orc_grid *g;
  int nb, nzb, nxb, nxb1;
  double *xb, *yb, *zb, *normals, *sb, *ub;
  /* stencils s = 0 u, 1 v, 2 w, 3 c: [s][l*NW + count] */
  int *xi[4], *yi[4], *zi[4];
  double *w[4], *dxn[4];
  double *E1nHc, *E1nH;       /* Heaviside coefficients (nb, 3nb) */
  double *Hu_int, *Hv_int, *Hw_int; /* interior masks (field shapes) */
};

/* deltafnc: src/immersed_boundary_operators.f90:270-294 */
double orc_deltafnc(double x1, double x2, double dr) {
  const double pi = 4.0 * atan(1.0);
  double r = fabs(x1 - x2);
  double r1 = r / dr;
  double r2 = r1 * r1;
  double d;
  if (r1 <= 1.0) {
    d = 17.0 / 48.0 + sqrt(3.0) * pi / 108.0 + r1 / 4.0 - r2 / 4.0 +
        (1.0 - 2.0 * r1) / 16.0 * sqrt(-12.0 * r2 + 12.0 * r1 + 1.0) -
        sqrt(3.0) / 12.0 * asin(sqrt(3.0) / 2.0 * (2.0 * r1 - 1.0));
  } else if (r1 <= 2.0) {
    d = 55.0 / 48.0 - sqrt(3.0) * pi / 108 - 13.0 * r1 / 12.0 + r2 / 4.0 +
        (2.0 * r1 - 3.0) / 48.0 * sqrt(-12.0 * r2 + 36.0 * r1 - 23.0) +
        sqrt(3.0) / 36.0 * asin(sqrt(3.0) / 2.0 * (2.0 * r1 - 3.0));
  } else {
    d = 0.0;
  }
  return d / dr;
}

static double *dup(const double *a, size_t n) {
  double *b = malloc(n * sizeof(double));
  memcpy(b, a, n * sizeof(double));
  return b;
}

static long floordiv_d(double a) { return (long)floor(a); }

/* setup_IB_operators: src/immersed_boundary_operators.f90:19-256 (one domain).
 * yref[l] is y_ref_index(l), the face whose coordinate anchors the y pivots
 * (NULL: 1 for every marker). */
orc_ib *orc_ib_new(const orc_grid *g, int nb, const double *xb, const double *yb,
                   const double *zb, const double *normals, const double *sb,
                   const double *ub, const int *yref_in, int nzb, int nxb, int nxb1) {
  orc_ib *ib = calloc(1, sizeof(orc_ib));
  ib->g = g;
  ib->nb = nb; ib->nzb = nzb; ib->nxb = nxb; ib->nxb1 = nxb1;
  ib->xb = dup(xb, nb); ib->yb = dup(yb, nb); ib->zb = dup(zb, nb);
  ib->normals = dup(normals, 3 * (size_t)nb);
  ib->sb = dup(sb, nb);
  ib->ub = dup(ub, 3 * (size_t)nb);
  for (int s = 0; s < 4; s++) {
    ib->xi[s] = malloc(sizeof(int) * NW * nb);
    ib->yi[s] = malloc(sizeof(int) * NW * nb);
    ib->zi[s] = malloc(sizeof(int) * NW * nb);
    ib->w[s] = malloc(sizeof(double) * NW * nb);
    ib->dxn[s] = malloc(sizeof(double) * NW * nb);
  }
  const int nx = g->nx, nz = g->nz, nzg = g->nzg;
  const int nxm = nx - 1, nzm = nz - 1;
  const double dx = g->dx, dz = g->dz, dymin = g->dymin, Lxp = g->Lxp, Lzp = g->Lzp;
  const double *x = g->x, *y = g->y, *z = g->z;
  /* xm(i) = (x(i)+x(i+1))/2 = xg(i+1) */
  for (int l = 0; l < nb; l++) {
    const int yref = yref_in ? yref_in[l] : 1;
    /* pivots (:28-42); 0.5*dx uses the real(4) literal 0.5 (exact) */
    int x_piv = (int)floordiv_d(xb[l] / dx) + 1;
    int xm_piv = (int)floordiv_d((xb[l] + 0.5 * dx) / dx);
    int y_piv = (int)floordiv_d((yb[l] - y[yref - 1]) / dymin) + yref;
    int ym_piv = (int)floordiv_d((yb[l] - (y[yref - 1] + 0.5 * dymin)) / dymin) + yref;
    int z_piv = (int)floordiv_d(zb[l] / dz) + 1;
    int zm_piv = (int)floordiv_d((zb[l] + 0.5 * dz) / dz) + 1;
    const double n0 = normals[l], n1 = normals[nb + l], n2 = normals[2 * nb + l];
    for (int s = 0; s < 4; s++) {
      int count = 0;
      for (int k = -2; k <= 2; k++)
        for (int j = -2; j <= 2; j++)
          for (int i = -2; i <= 2; i++) {
            int ii, jj, kk, xs, zs, iip, kkp, X, Y, Zi;
            double xgd, ygd, zgd;
            if (s == 0) { /* u (:46-98) */
              ii = x_piv + i; jj = ym_piv + j; kk = zm_piv + k;
              xs = (int)floor((double)(ii - 1) / (double)(nx - 2));
              iip = ii - xs * (nx - 2);
              zs = (int)floor((double)(kk - 2) / (double)nzm);
              kkp = kk - zs * (nzm - 1);
              X = (iip == 1) ? nx - 1 : iip;
              Y = jj + 1;
              Zi = kkp;
              xgd = x[iip - 1] + xs * Lxp;
              ygd = 0.5 * (y[jj - 1] + y[jj]);             /* ym_global(jj) */
              zgd = 0.5 * (z[kkp - 2] + z[kkp - 1]) + zs * Lzp; /* zm_global(kkp-1) */
            } else if (s == 1) { /* v (:101-148) */
              ii = xm_piv + i; jj = y_piv + j; kk = zm_piv + k;
              xs = (int)floor((double)(ii - 1) / (double)(nxm - 1));
              iip = ii - xs * (nxm - 1);
              zs = (int)floor((double)(kk - 2) / (double)nzm);
              kkp = kk - zs * (nzm - 1);
              X = iip + 1;
              Y = jj;
              Zi = kkp;
              xgd = 0.5 * (x[iip - 1] + x[iip]) + xs * Lxp;
              ygd = y[jj - 1];
              zgd = 0.5 * (z[kkp - 2] + z[kkp - 1]) + zs * Lzp;
            } else if (s == 2) { /* w (:151-194), x offset +i-1 as in the reference */
              ii = xm_piv + i - 1; jj = ym_piv + j; kk = z_piv + k;
              xs = (int)floor((double)(ii - 1) / (double)(nxm - 1));
              iip = ii - xs * (nxm - 1);
              zs = (int)floor((double)(kk - 2) / (double)(nz - 2));
              kkp = kk - zs * (nz - 2);
              X = iip + 1;
              Y = jj + 1;
              Zi = kkp;
              xgd = 0.5 * (x[iip - 1] + x[iip]) + xs * Lxp;
              ygd = 0.5 * (y[jj - 1] + y[jj]);
              zgd = z[kkp - 1] + zs * Lzp;
            } else { /* c (:197-244) */
              ii = xm_piv + i; jj = ym_piv + j; kk = zm_piv + k;
              xs = (int)floor((double)(ii - 1) / (double)(nxm - 1));
              iip = ii - xs * (nxm - 1);
              zs = (int)floor((double)(kk - 2) / (double)nzm);
              kkp = kk - zs * (nzm - 1);
              X = iip + 1;
              Y = jj + 1;
              Zi = kkp;
              xgd = 0.5 * (x[iip - 1] + x[iip]) + xs * Lxp;
              ygd = 0.5 * (y[jj - 1] + y[jj]);
              zgd = 0.5 * (z[kkp - 2] + z[kkp - 1]) + zs * Lzp;
            }
            /* canonical periodic plane for centre z (nzg-1 == 2) */
            if (s != 2 && Zi == nzg - 1) Zi = 2;
            size_t o = (size_t)l * NW + count;
            ib->xi[s][o] = X;
            ib->yi[s][o] = Y;
            ib->zi[s][o] = Zi;
            ib->w[s][o] = dx * dymin * dz * orc_deltafnc(xgd, xb[l], dx) *
                          orc_deltafnc(ygd, yb[l], dymin) * orc_deltafnc(zgd, zb[l], dz);
            ib->dxn[s][o] = n0 * (xgd - xb[l]) + n1 * (ygd - yb[l]) + n2 * (zgd - zb[l]);
            count++;
          }
    }
  }
  return ib;
}

void orc_ib_free(orc_ib *ib) {
  if (!ib) return;
  free(ib->xb); free(ib->yb); free(ib->zb); free(ib->normals); free(ib->sb); free(ib->ub);
  for (int s = 0; s < 4; s++) {
    free(ib->xi[s]); free(ib->yi[s]); free(ib->zi[s]); free(ib->w[s]); free(ib->dxn[s]);
  }
  free(ib->E1nHc); free(ib->E1nH); free(ib->Hu_int); free(ib->Hv_int); free(ib->Hw_int);
  free(ib);
}

/* copy out stencil s (0 u, 1 v, 2 w, 3 c): xi, yi, zi [nb*125], w, dxn */
void orc_ib_get_stencil(const orc_ib *ib, int s, int *xi, int *yi, int *zi, double *w,
                        double *dxn) {
  size_t n = (size_t)ib->nb * NW;
  memcpy(xi, ib->xi[s], n * sizeof(int));
  memcpy(yi, ib->yi[s], n * sizeof(int));
  memcpy(zi, ib->zi[s], n * sizeof(int));
  memcpy(w, ib->w[s], n * sizeof(double));
  memcpy(dxn, ib->dxn[s], n * sizeof(double));
}

/* field extents per stencil */
static void dims(const orc_grid *g, int s, int *n1, int *n2) {
  if (s == 0) { *n1 = g->nx; *n2 = g->nyg; }
  else if (s == 1) { *n1 = g->nxg; *n2 = g->ny; }
  else { *n1 = g->nxg; *n2 = g->nyg; }
}

/* local_regT (E, :409-454) and local_regT_1n (E1n, :456-501):
 * out[c*nb + l] = sum_i w * F(idx) [* dxn] for c = u, v, w. */
void orc_ib_regT(const orc_ib *ib, const double *U, const double *V, const double *W,
                 int with_dxn, double *out) {
  const orc_grid *g = ib->g;
  const double *F[3] = {U, V, W};
  int nb = ib->nb;
  for (int c = 0; c < 3; c++) {
    int n1, n2;
    dims(g, c, &n1, &n2);
    for (int l = 0; l < nb; l++) {
      double acc = 0.0;
      for (int i = 0; i < NW; i++) {
        size_t o = (size_t)l * NW + i;
        double v = ib->w[c][o] * F[c][IX(ib->xi[c][o], ib->yi[c][o], ib->zi[c][o], n1, n2)];
        if (with_dxn) v = v * ib->dxn[c][o];
        acc = acc + v;
      }
      out[(size_t)c * nb + l] = acc;
    }
  }
}

/* local_regTc_1n (:503-526): sum_i c_w * P(idx) * dxnc, P in rhs_p shape */
void orc_ib_regTc_1n(const orc_ib *ib, const double *P, double *out) {
  const orc_grid *g = ib->g;
  for (int l = 0; l < ib->nb; l++) {
    double acc = 0.0;
    for (int i = 0; i < NW; i++) {
      size_t o = (size_t)l * NW + i;
      acc = acc + ib->w[3][o] * P[RX(ib->xi[3][o], ib->yi[3][o], ib->zi[3][o], g)] * ib->dxn[3][o];
    }
    out[l] = acc;
  }
}

/* regu / regv / regw with local_reg (:359-400, :544-603): F = 0 then
 * F(idx) += w * (1/(dx*dymin*dz)) * sb(l) * f(l), marker-major order. */
void orc_ib_reg(const orc_ib *ib, int comp, const double *f, double *F) {
  const orc_grid *g = ib->g;
  int n1, n2;
  dims(g, comp, &n1, &n2);
  size_t nf = (size_t)n1 * n2 * (comp == 2 ? g->nz : g->nzg);
  memset(F, 0, nf * sizeof(double));
  const double factor = 1.0 / (g->dx * g->dymin * g->dz);
  for (int l = 0; l < ib->nb; l++)
    for (int i = 0; i < NW; i++) {
      size_t o = (size_t)l * NW + i;
      size_t p = IX(ib->xi[comp][o], ib->yi[comp][o], ib->zi[comp][o], n1, n2);
      F[p] = F[p] + ib->w[comp][o] * factor * ib->sb[l] * f[l];
    }
}

/* remove_mean_per_body for the two-group cylinder layout
 * (src/immersed_boundary_geometry.f90:400-442): per z slab j the markers
 * (j-1)*nxb + 1..nxb1 form body 1 and the rest body 2; per-slab sums are
 * accumulated slab by slab. */
void orc_ib_remove_mean(const orc_ib *ib, double *f) {
  double avg[2] = {0.0, 0.0};
  const int nxb = ib->nxb, nxb1 = ib->nxb1, nxb2 = nxb - nxb1;
  for (int j = 0; j < ib->nzb; j++) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < nxb1; i++) s1 += f[(size_t)j * nxb + i];
    for (int i = nxb1; i < nxb; i++) s2 += f[(size_t)j * nxb + i];
    avg[0] = avg[0] + s1;
    avg[1] = avg[1] + s2;
  }
  avg[0] = avg[0] / (double)(nxb1 * ib->nzb);
  avg[1] = avg[1] / (double)(nxb2 * ib->nzb);
  for (int j = 0; j < ib->nzb; j++) {
    for (int i = 0; i < nxb1; i++) f[(size_t)j * nxb + i] -= avg[0];
    for (int i = nxb1; i < nxb; i++) f[(size_t)j * nxb + i] -= avg[1];
  }
}

/* compute_heaviside: src/heaviside.f90:20-54.  Stores E1nHc_exterior,
 * E1nH_exterior and the interior masks in ib; optional outputs (NULL to
 * skip): Hc_ext (rhs_p shape), Hu/Hv/Hw_ext (field shapes). */
void orc_ib_heaviside(orc_ib *ib, double *Hc_out, double *Hu_out, double *Hv_out,
                      double *Hw_out) {
  const orc_grid *g = ib->g;
  const int nx = g->nx, ny = g->ny, nz = g->nz, nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
  size_t nu = (size_t)nx * nyg * nzg, nv = (size_t)nxg * ny * nzg, nw = (size_t)nxg * nyg * nz;
  size_t nr = (size_t)(nxg - 2) * (nyg - 2) * (nzg - 1);
  int nb = ib->nb;
  double *Ur = malloc(nu * 8), *Vr = malloc(nv * 8), *Wr = malloc(nw * 8);
  double *Hc = malloc(nr * 8);
  orc_ib_reg(ib, 0, ib->normals, Ur);
  orc_ib_reg(ib, 1, ib->normals + nb, Vr);
  orc_ib_reg(ib, 2, ib->normals + 2 * (size_t)nb, Wr);
  orc_apply_bc(g, Ur, Vr, Wr);
  orc_divergence(g, Hc, Ur, Vr, Wr);
  for (size_t q = 0; q < nr; q++) Hc[q] = -Hc[q];
  orc_poisson(g, Hc, 0, 1.0, NULL, 0, 0); /* rk_step is not 1 at set-up */
  for (size_t q = 0; q < nr; q++) Hc[q] = 1.0 - Hc[q]; /* Hc_exterior */
  double *Hu = calloc(nu, 8), *Hv = calloc(nv, 8), *Hw = calloc(nw, 8);
  /* interpolate_x(Hc_ext, Hu_ext(2:nx-1, 2:nyg-1, 2:nzg)) and y, z */
  for (int k = 2; k <= nzg; k++)
    for (int j = 2; j <= nyg - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        Hu[IX(i, j, k, nx, nyg)] = 0.5 * (Hc[RX(i, j, k, g)] + Hc[RX(i + 1, j, k, g)]);
  for (int k = 2; k <= nzg; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nxg - 1; i++)
        Hv[IX(i, j, k, nxg, ny)] = 0.5 * (Hc[RX(i, j, k, g)] + Hc[RX(i, j + 1, k, g)]);
  for (int k = 2; k <= nz; k++)
    for (int j = 2; j <= nyg - 1; j++)
      for (int i = 2; i <= nxg - 1; i++)
        Hw[IX(i, j, k, nxg, nyg)] = 0.5 * (Hc[RX(i, j, k, g)] + Hc[RX(i, j, k + 1, g)]);
  orc_apply_bc(g, Hu, Hv, Hw);
  free(ib->Hu_int); free(ib->Hv_int); free(ib->Hw_int);
  ib->Hu_int = malloc(nu * 8); ib->Hv_int = malloc(nv * 8); ib->Hw_int = malloc(nw * 8);
  for (size_t q = 0; q < nu; q++) ib->Hu_int[q] = 1.0 - Hu[q];
  for (size_t q = 0; q < nv; q++) ib->Hv_int[q] = 1.0 - Hv[q];
  for (size_t q = 0; q < nw; q++) ib->Hw_int[q] = 1.0 - Hw[q];
  orc_apply_bc(g, ib->Hu_int, ib->Hv_int, ib->Hw_int);
  free(ib->E1nHc); free(ib->E1nH);
  ib->E1nHc = malloc(nb * 8);
  ib->E1nH = malloc(3 * (size_t)nb * 8);
  orc_ib_regTc_1n(ib, Hc, ib->E1nHc);
  orc_ib_regT(ib, Hu, Hv, Hw, 1, ib->E1nH);
  if (Hc_out) memcpy(Hc_out, Hc, nr * 8);
  if (Hu_out) memcpy(Hu_out, Hu, nu * 8);
  if (Hv_out) memcpy(Hv_out, Hv, nv * 8);
  if (Hw_out) memcpy(Hw_out, Hw, nw * 8);
  free(Ur); free(Vr); free(Wr); free(Hc); free(Hu); free(Hv); free(Hw);
}

void orc_ib_get_heaviside(const orc_ib *ib, double *E1nHc, double *E1nH, double *Hu_int,
                          double *Hv_int, double *Hw_int) {
  const orc_grid *g = ib->g;
  size_t nu = (size_t)g->nx * g->nyg * g->nzg, nv = (size_t)g->nxg * g->ny * g->nzg,
         nw = (size_t)g->nxg * g->nyg * g->nz;
  if (E1nHc) memcpy(E1nHc, ib->E1nHc, ib->nb * 8);
  if (E1nH) memcpy(E1nH, ib->E1nH, 3 * (size_t)ib->nb * 8);
  if (Hu_int) memcpy(Hu_int, ib->Hu_int, nu * 8);
  if (Hv_int) memcpy(Hv_int, ib->Hv_int, nv * 8);
  if (Hw_int) memcpy(Hw_int, ib->Hw_int, nw * 8);
}

/* schur: src/projection.f90:98-144.  f = (dudn[3nb], p_jump[nb]) -> Sf.
 * U, V, W are scratch (the reference overwrites the global velocity). */
typedef struct {
  double *Ureg, *Vreg, *Wreg, *U, *V, *W, *rhs, *aux;
} schur_ws;

static void schur_ws_alloc(const orc_grid *g, int nb, schur_ws *s) {
  size_t nu = (size_t)g->nx * g->nyg * g->nzg, nv = (size_t)g->nxg * g->ny * g->nzg,
         nw = (size_t)g->nxg * g->nyg * g->nz;
  size_t nr = (size_t)(g->nxg - 2) * (g->nyg - 2) * (g->nzg - 1);
  s->Ureg = malloc(nu * 8); s->Vreg = malloc(nv * 8); s->Wreg = malloc(nw * 8);
  s->U = malloc(nu * 8); s->V = malloc(nv * 8); s->W = malloc(nw * 8);
  s->rhs = malloc(nr * 8);
  s->aux = malloc(4 * (size_t)nb * 8);
}
static void schur_ws_free(schur_ws *s) {
  free(s->Ureg); free(s->Vreg); free(s->Wreg); free(s->U); free(s->V); free(s->W);
  free(s->rhs); free(s->aux);
}

static void schur_apply(const orc_ib *ib, schur_ws *s, const double *f, double *Sf, double dt,
                        double nu, int rk_step, double *P, int nthreads) {
  const orc_grid *g = ib->g;
  const int nb = ib->nb;
  const double *dudn = f, *pj = f + 3 * (size_t)nb;
  size_t nu_ = (size_t)g->nx * g->nyg * g->nzg, nv = (size_t)g->nxg * g->ny * g->nzg,
         nw = (size_t)g->nxg * g->nyg * g->nz;
  size_t nr = (size_t)(g->nxg - 2) * (g->nyg - 2) * (g->nzg - 1);
  double *Freg[3] = {s->Ureg, s->Vreg, s->Wreg};
  for (int c = 0; c < 3; c++) { /* :109-114 */
    for (int l = 0; l < nb; l++)
      s->aux[l] = dt * nu * dudn[(size_t)c * nb + l] - dt * pj[l] * ib->normals[(size_t)c * nb + l];
    orc_ib_reg(ib, c, s->aux, Freg[c]);
  }
  orc_apply_bc(g, s->Ureg, s->Vreg, s->Wreg);               /* :116 */
  orc_divergence(g, s->rhs, s->Ureg, s->Vreg, s->Wreg);     /* :119 */
  orc_poisson(g, s->rhs, rk_step, dt, P, 0, nthreads);      /* :122 */
  for (size_t q = 0; q < nr; q++) s->rhs[q] = s->rhs[q] / dt; /* :123 */
  double *E1np = s->aux;
  orc_ib_regTc_1n(ib, s->rhs, E1np);                        /* :125 */
  for (int l = 0; l < nb; l++) Sf[3 * (size_t)nb + l] = -E1np[l];
  orc_ib_remove_mean(ib, Sf + 3 * (size_t)nb);              /* :127 */
  for (int l = 0; l < nb; l++)
    Sf[3 * (size_t)nb + l] = Sf[3 * (size_t)nb + l] - ib->E1nHc[l] * pj[l];
  for (size_t q = 0; q < nr; q++) s->rhs[q] = dt * s->rhs[q]; /* :131 */
  orc_gradient(g, s->U, s->V, s->W, s->rhs);                /* :132 */
  orc_apply_bc(g, s->U, s->V, s->W);                        /* :133 */
  for (size_t q = 0; q < nu_; q++) s->U[q] = s->U[q] - s->Ureg[q]; /* :136-138 */
  for (size_t q = 0; q < nv; q++) s->V[q] = s->V[q] - s->Vreg[q];
  for (size_t q = 0; q < nw; q++) s->W[q] = s->W[q] - s->Wreg[q];
  orc_ib_regT(ib, s->U, s->V, s->W, 0, Sf);                 /* :141 */
  for (size_t q = 0; q < 3 * (size_t)nb; q++) Sf[q] = Sf[q] - ib->E1nH[q] * dudn[q];
}

void orc_ib_schur(const orc_ib *ib, const double *f, double *Sf, double dt, double nu,
                  int nthreads) {
  schur_ws s;
  schur_ws_alloc(ib->g, ib->nb, &s);
  schur_apply(ib, &s, f, Sf, dt, nu, 0, NULL, nthreads);
  schur_ws_free(&s);
}

static double dotp(const double *a, const double *b, size_t n) {
  double s = 0.0;
  for (size_t q = 0; q < n; q++) s = s + a[q] * b[q];
  return s;
}

/* bicgstab: src/projection.f90:146-186 (no preconditioner; absolute test
 * ||r||^2 < tol^2; at least one iteration).  Returns the iteration count,
 * *err receives the last ||r||^2. */
int orc_ib_bicgstab(const orc_ib *ib, double *x, const double *b, double tol, int maxit,
                    double dt, double nu, int rk_step, double *P, double *err, int nthreads) {
  const size_t n = 4 * (size_t)ib->nb;
  schur_ws s;
  schur_ws_alloc(ib->g, ib->nb, &s);
  double *r = malloc(n * 8), *rhat = malloc(n * 8), *p = calloc(n, 8), *v = calloc(n, 8);
  double *h = malloc(n * 8), *sv = malloc(n * 8), *tv = malloc(n * 8);
  double error = 1.0, eps = tol * tol;
  int iter = 0;
  schur_apply(ib, &s, x, r, dt, nu, rk_step, P, nthreads);
  for (size_t q = 0; q < n; q++) r[q] = b[q] - r[q];
  memcpy(rhat, r, n * 8);
  double rho_o = 1.0, alpha = 1.0, om = 1.0;
  while (iter < maxit && error >= eps) {
    double rho_n = dotp(rhat, r, n);
    double bta = (rho_n / rho_o) * (alpha / om);
    rho_o = rho_n;
    for (size_t q = 0; q < n; q++) p[q] = r[q] + bta * (p[q] - om * v[q]);
    schur_apply(ib, &s, p, v, dt, nu, rk_step, P, nthreads);
    alpha = rho_n / dotp(rhat, v, n);
    for (size_t q = 0; q < n; q++) h[q] = x[q] + alpha * p[q];
    for (size_t q = 0; q < n; q++) sv[q] = r[q] - alpha * v[q];
    schur_apply(ib, &s, sv, tv, dt, nu, rk_step, P, nthreads);
    om = dotp(tv, sv, n) / dotp(tv, tv, n);
    for (size_t q = 0; q < n; q++) x[q] = h[q] + om * sv[q];
    for (size_t q = 0; q < n; q++) r[q] = sv[q] - om * tv[q];
    error = dotp(r, r, n);
    iter = iter + 1;
  }
  if (err) *err = error;
  free(r); free(rhat); free(p); free(v); free(h); free(sv); free(tv);
  schur_ws_free(&s);
  return iter;
}

/* compute_IB_projection: src/projection.f90:48-96, after a non-IB projection
 * (Uint, Vint, Wint its interim copies, Pint its P_interim).  fb (4nb) is
 * the BiCGSTAB initial guess on input and the force on output. */
int orc_ib_projection(const orc_ib *ib, double *U, double *V, double *W, const double *Uint,
                      const double *Vint, const double *Wint, double *Pint, double *P,
                      double *fb, double dt, double nu, double tol, int maxit, int rk_step,
                      double *err, int nthreads) {
  const orc_grid *g = ib->g;
  const int nb = ib->nb;
  size_t nu_ = (size_t)g->nx * g->nyg * g->nzg, nv = (size_t)g->nxg * g->ny * g->nzg,
         nw = (size_t)g->nxg * g->nyg * g->nz;
  size_t nr = (size_t)(g->nxg - 2) * (g->nyg - 2) * (g->nzg - 1);
  double *rhs_ib = malloc(4 * (size_t)nb * 8);
  orc_ib_regT(ib, U, V, W, 0, rhs_ib);                              /* :51 */
  for (size_t q = 0; q < 3 * (size_t)nb; q++) rhs_ib[q] = -rhs_ib[q] + ib->ub[q]; /* :52 */
  orc_ib_regTc_1n(ib, Pint, rhs_ib + 3 * (size_t)nb);               /* :53 */
  for (int l = 0; l < nb; l++) rhs_ib[3 * (size_t)nb + l] = -rhs_ib[3 * (size_t)nb + l];
  orc_ib_remove_mean(ib, rhs_ib + 3 * (size_t)nb);                  /* :57 */
  int iters = orc_ib_bicgstab(ib, fb, rhs_ib, tol, maxit, dt, nu, rk_step, P, err, nthreads);
  schur_ws s;
  schur_ws_alloc(g, nb, &s);
  const double *dudn = fb, *pj = fb + 3 * (size_t)nb;
  double *Freg[3] = {s.Ureg, s.Vreg, s.Wreg};
  for (int c = 0; c < 3; c++) {                                     /* :66-71 */
    for (int l = 0; l < nb; l++)
      s.aux[l] = dt * nu * dudn[(size_t)c * nb + l] - dt * pj[l] * ib->normals[(size_t)c * nb + l];
    orc_ib_reg(ib, c, s.aux, Freg[c]);
  }
  orc_apply_bc(g, s.Ureg, s.Vreg, s.Wreg);                          /* :73 */
  orc_divergence(g, s.rhs, s.Ureg, s.Vreg, s.Wreg);                 /* :76 */
  orc_poisson(g, s.rhs, rk_step, dt, P, 0, nthreads);               /* :79 */
  for (size_t q = 0; q < nr; q++) s.rhs[q] = s.rhs[q] / dt;         /* :80 */
  for (size_t q = 0; q < nr; q++) s.rhs[q] = Pint[q] - s.rhs[q];    /* :83 */
  for (size_t q = 0; q < nr; q++) s.rhs[q] = dt * s.rhs[q];         /* :86 */
  orc_gradient(g, U, V, W, s.rhs);                                  /* :87 */
  for (size_t q = 0; q < nu_; q++) U[q] = Uint[q] - s.Ureg[q] - U[q]; /* :90-92 */
  for (size_t q = 0; q < nv; q++) V[q] = Vint[q] - s.Vreg[q] - V[q];
  for (size_t q = 0; q < nw; q++) W[q] = Wint[q] - s.Wreg[q] - W[q];
  orc_apply_bc(g, U, V, W);                                         /* :94 */
  schur_ws_free(&s);
  free(rhs_ib);
  return iters;
}

/* check_slip: src/monitor.f90:223-231 with dudn = fb(1:3nb) */
double orc_ib_max_slip(const orc_ib *ib, const double *U, const double *V, const double *W,
                       const double *fb) {
  const int nb = ib->nb;
  double *Eu = malloc(3 * (size_t)nb * 8);
  orc_ib_regT(ib, U, V, W, 0, Eu);
  double m = 0.0;
  for (size_t q = 0; q < 3 * (size_t)nb; q++) {
    double a = fabs(Eu[q] - ib->E1nH[q] * fb[q] - ib->ub[q]);
    if (a > m) m = a;
  }
  free(Eu);
  return m;
}
