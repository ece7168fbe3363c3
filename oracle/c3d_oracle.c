/*
 * c3d_oracle.c -- CPU FP64 restatement of the channel3D fractional-step
 * projection path (grid, BCs, staggered div/grad, FFT-xz + tridiagonal-y
 * Poisson solve, non-IB projection).
 *
 * TEST INFRASTRUCTURE ONLY: the reference against which the HIP path is
 * tested, and the CPU baseline timed beside it.  Never linked into the
 * product library.
 *
 * Every routine cites the reference lines it restates (paths relative to the
 * reference repository root).  Index conventions are documented in
 * c3d_oracle.h.
 */
#include "c3d_oracle.h"
#include "c3d_oracle_impl.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* complex helpers (explicit arithmetic: no libgcc __muldc3 semantics)       */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
/* Complex division.  For a real-valued denominator (every use in the
 * Poisson path: Dyy, kxx, kzz are real) this is the componentwise quotient,
 * which is what any complex division algorithm returns for im(b) == 0.
 * Otherwise Smith's algorithm. */
static inline cplx cdiv(cplx a, cplx b) {
  cplx r;
  if (b.im == 0.0) {
    r.re = a.re / b.re;
    r.im = a.im / b.re;
    return r;
  }
  if (fabs(b.re) >= fabs(b.im)) {
    double t = b.im / b.re, den = b.re + b.im * t;
    r.re = (a.re + a.im * t) / den;
    r.im = (a.im - a.re * t) / den;
  } else {
    double t = b.re / b.im, den = b.re * t + b.im;
    r.re = (a.re * t + a.im) / den;
    r.im = (a.im * t - a.re) / den;
  }
  return r;
}
static inline double cabs1(cplx a) { return fabs(a.re) + fabs(a.im); }

/* ------------------------------------------------------------------------ */
/* grid                                                                     */
/* ------------------------------------------------------------------------ */


static double min_diff(const double *a, int n) {
  double m = a[1] - a[0];
  for (int i = 1; i < n - 1; i++) {
    double d = a[i + 1] - a[i];
    if (d < m) m = d;
  }
  return m;
}

/* Centres with mirror ghosts: src/initialization.f90:173-199. */
static void centres(const double *f, int n, double *g) {
  /* f: faces 1..n ; g: 1..n+1 with g(2..n) = midpoints */
  for (int i = 0; i < n - 1; i++) g[i + 1] = 0.5 * (f[i] + f[i + 1]);
  g[0] = g[1] - 2.0 * (g[1] - f[0]);
  g[n] = g[n - 1] + 2.0 * (f[n - 1] - g[n - 1]);
}

/* Modified wavenumber 2(cos(2 pi i/n) - 1)/d^2: src/initialization.f90:290-305.
 * The reference writes dx**2d0; it is evaluated here as dx*dx. */
static void modified_wavenumbers(int n, double d, double *k) {
  const double pi = 4.0 * atan(1.0); /* src/global.f90:34 */
  int half = (n + 1) / 2;            /* Ceiling(n/2) */
  for (int i = 0; i < n; i++) k[i] = 0.0;
  for (int i = 0; i <= half && i < n; i++)
    k[i] = 2.0 * (cos(2.0 * pi * (double)i / (double)n) - 1.0) / (d * d);
  for (int i = half + 1; i < n; i++)
    k[i] = 2.0 * (cos(2.0 * pi * (double)(-n + i) / (double)n) - 1.0) / (d * d);
}

/* create_stretched_grid, src/input_output.f90:208-247: tanh points on
 * [-1, 1] over the middle n_total - 2(n_uniform-1) points, scaled so that the
 * stretched part plus two half buffers span L, extended by n_uniform - 1
 * points of the smallest spacing at each end. */
static void stretched_with_buffers(double *grid, double L, int n_total, int n_uniform,
                                   double alpha) {
  const int nun = n_total - 2 * (n_uniform - 1);
  double *gu = malloc(sizeof(double) * nun);
  for (int i = 0; i < nun; i++) gu[i] = (double)i;
  double mx = gu[0];
  for (int i = 1; i < nun; i++) if (gu[i] > mx) mx = gu[i];
  for (int i = 0; i < nun; i++) gu[i] = 2.0 * gu[i] / mx - 1.0;
  if (alpha > 0.0)
    for (int i = 0; i < nun; i++) gu[i] = tanh(alpha * gu[i]) / tanh(alpha);
  const double ds_un = gu[1] - gu[0];
  const double ds_sc = L * ds_un / (2.0 + (double)(n_uniform - 1) * ds_un);
  for (int i = 0; i < nun; i++) grid[n_uniform - 1 + i] = gu[i] * ds_sc / ds_un;
  for (int i = 2; i <= n_uniform; i++) {
    grid[n_total - n_uniform + i - 1] = grid[n_total - n_uniform + i - 2] + ds_sc;
    grid[n_uniform - i] = grid[n_uniform - i + 1] - ds_sc;
  }
  free(gu);
}

orc_grid *orc_grid_new(int nx, int ny, int nz, double Lxp, double Lzp,
                       double Ly, double alpha, int grid_type, double min_buffer_width) {
  /* src/initialization.f90:46-47 (even sizes), one domain */
  if (nx < 6 || ny < 4 || nz < 6 || (nx % 2) || (nz % 2)) return NULL;
  if (grid_type < 0 || grid_type > 2) return NULL;
  orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
  g->nx = nx; g->ny = ny; g->nz = nz; g->n_uniform = 1;
  g->nxg = nx + 1; g->nyg = ny + 1; g->nzg = nz + 1;
  g->nxp = nx - 2; g->nzp = nz - 2; g->nym = ny - 1;
  g->x = calloc(nx, sizeof(double)); g->y = calloc(ny, sizeof(double));
  g->z = calloc(nz, sizeof(double));
  g->xg = calloc(nx + 1, sizeof(double)); g->yg = calloc(ny + 1, sizeof(double));
  g->zg = calloc(nz + 1, sizeof(double));
  g->kxx = calloc(g->nxp, sizeof(double)); g->kzz = calloc(g->nzp, sizeof(double));
  g->dyy_d = calloc(g->nym, sizeof(double));
  g->dyy_l = calloc(g->nym, sizeof(double));
  g->dyy_u = calloc(g->nym, sizeof(double));

  /* create_grid: src/input_output.f90:146-154 */
  for (int i = 0; i < nx; i++) g->x[i] = Lxp * (double)i / (double)(nx - 2);
  for (int i = 0; i < nz; i++) g->z[i] = Lzp * (double)i / (double)(nz - 2);
  if (grid_type == 0) { /* :157-162 */
    for (int i = 0; i < ny; i++) g->y[i] = Ly * (double)i / (double)(ny - 1);
  } else if (grid_type == 2) { /* :180-203, suppy = 2 */
    int nu = 1;
    for (;;) {
      stretched_with_buffers(g->y, Ly, ny, nu, alpha);
      const double dy1 = g->y[1] - g->y[0];
      if ((double)nu * dy1 >= min_buffer_width + (2 * 2 + 2) * dy1) break;
      nu++;
      if (2 * nu >= ny) { orc_grid_free(g); return NULL; }
    }
    g->n_uniform = nu;
    for (int i = 0; i < ny; i++) g->y[i] = g->y[i] + 1.0;
  } else { /* :164-178 */
    for (int i = 0; i < ny; i++) g->y[i] = Ly * (double)i / (double)(ny - 1) - 1.0;
    if (alpha > 0.0)
      for (int i = 0; i < ny; i++) g->y[i] = tanh(alpha * g->y[i]) / tanh(alpha);
    double mn = g->y[0];
    for (int i = 1; i < ny; i++) if (g->y[i] < mn) mn = g->y[i];
    for (int i = 0; i < ny; i++) g->y[i] = g->y[i] - mn;
    double mx = g->y[0];
    for (int i = 1; i < ny; i++) if (g->y[i] > mx) mx = g->y[i];
    for (int i = 0; i < ny; i++) g->y[i] = g->y[i] * Ly / mx;
  }
  centres(g->x, nx, g->xg);
  centres(g->y, ny, g->yg);
  centres(g->z, nz, g->zg);
  /* src/initialization.f90:212-219,247-252 */
  g->dxmin = min_diff(g->xg, nx + 1);
  g->dymin = min_diff(g->yg, ny + 1);
  g->dzmin = min_diff(g->zg, nz + 1);
  g->Lx = g->x[nx - 1] - g->x[0];
  g->Ly = g->y[ny - 1] - g->y[0];
  g->Lz = g->z[nz - 1] - g->z[0];
  g->dx = g->dxmin;
  g->dz = g->dzmin;
  g->Lxp = g->Lx - g->dx;
  g->Lzp = g->Lz - g->dz;
  modified_wavenumbers(g->nxp, g->dx, g->kxx);
  modified_wavenumbers(g->nzp, g->dz, g->kzz);

  /* Dyy: src/initialization.f90:353-383.  Fortran j (2..nyg-1) -> jj = j-2.
   * y(j) -> g->y[j-1], yg(j) -> g->yg[j-1]. */
  const double *y = g->y, *yg = g->yg;
  int nyg = g->nyg;
  for (int j = 2; j <= nyg - 1; j++) {
    double dyf = y[j - 1] - y[j - 2];
    double a = 1.0 / dyf / (yg[j] - yg[j - 1]);
    double b = 1.0 / dyf * (-1.0 / (yg[j] - yg[j - 1]) - 1.0 / (yg[j - 1] - yg[j - 2]));
    double c = 1.0 / dyf / (yg[j - 1] - yg[j - 2]);
    int jj = j - 2;
    if (j == 2) {
      g->dyy_d[jj] = b + c;
      g->dyy_u[jj] = a;
    } else if (j == nyg - 1) {
      g->dyy_d[jj] = a + b;
      g->dyy_l[jj - 1] = c; /* Dyy(nyg-1,nyg-2) */
    } else {
      g->dyy_d[jj] = b;
      g->dyy_u[jj] = a;     /* Dyy(j,j+1) */
      g->dyy_l[jj - 1] = c; /* Dyy(j,j-1) = DL(j-1) */
    }
  }
  return g;
}

void orc_grid_free(orc_grid *g) {
  if (!g) return;
  free(g->x); free(g->y); free(g->z); free(g->xg); free(g->yg); free(g->zg);
  free(g->kxx); free(g->kzz); free(g->dyy_d); free(g->dyy_l); free(g->dyy_u);
  free(g);
}

int orc_grid_get(const orc_grid *g, int what, double *out) {
  switch (what) {
  case 0: memcpy(out, g->x, g->nx * sizeof(double)); return g->nx;
  case 1: memcpy(out, g->y, g->ny * sizeof(double)); return g->ny;
  case 2: memcpy(out, g->z, g->nz * sizeof(double)); return g->nz;
  case 3: memcpy(out, g->xg, g->nxg * sizeof(double)); return g->nxg;
  case 4: memcpy(out, g->yg, g->nyg * sizeof(double)); return g->nyg;
  case 5: memcpy(out, g->zg, g->nzg * sizeof(double)); return g->nzg;
  case 6: memcpy(out, g->kxx, g->nxp * sizeof(double)); return g->nxp;
  case 7: memcpy(out, g->kzz, g->nzp * sizeof(double)); return g->nzp;
  case 8: memcpy(out, g->dyy_d, g->nym * sizeof(double)); return g->nym;
  case 9: memcpy(out, g->dyy_l, (g->nym - 1) * sizeof(double)); return g->nym - 1;
  case 10: memcpy(out, g->dyy_u, (g->nym - 1) * sizeof(double)); return g->nym - 1;
  case 11: {
    double s[10] = {g->dx, g->dz, g->dxmin, g->dymin, g->dzmin,
                    g->Lx, g->Ly, g->Lz, g->Lxp, g->Lzp};
    memcpy(out, s, sizeof(s));
    return 10;
  }
  case 12: out[0] = (double)g->n_uniform; return 1;
  default: return -1;
  }
}

/* ------------------------------------------------------------------------ */
/* index helpers                                                            */
/* ------------------------------------------------------------------------ */
#define IX(i, j, k, n1, n2) ((size_t)((i) - 1) + (size_t)(n1) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))
/* rhs_p(2:nxg-1, 2:nyg-1, 2:nzg): extents (nxg-2, nyg-2, nzg-1) */
#define RX(i, j, k, g) ((size_t)((i) - 2) + (size_t)((g)->nxg - 2) * ((size_t)((j) - 2) + (size_t)((g)->nyg - 2) * (size_t)((k) - 2)))

/* ------------------------------------------------------------------------ */
/* boundary conditions: src/boundary_conditions.f90:20-252 (one domain)      */
/* ------------------------------------------------------------------------ */
/* apply_periodic_bc_x (:67-86) */
static void periodic_x(double *F, int n1, int n2, int n3, int faces) {
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++) {
      if (faces) {
        F[IX(1, j, k, n1, n2)] = F[IX(n1 - 1, j, k, n1, n2)];
        F[IX(n1, j, k, n1, n2)] = F[IX(2, j, k, n1, n2)];
      } else {
        F[IX(1, j, k, n1, n2)] = F[IX(n1 - 2, j, k, n1, n2)];
        F[IX(n1 - 1, j, k, n1, n2)] = F[IX(2, j, k, n1, n2)];
        F[IX(n1, j, k, n1, n2)] = F[IX(3, j, k, n1, n2)];
      }
    }
}
/* apply_periodic_bc_z (:91-224) with myid == 0 == nprocs-1: all saves happen
 * before any plane is written. */
static void periodic_z(double *F, int n1, int n2, int n3, int faces) {
  size_t plane = (size_t)n1 * n2;
  if (faces) { /* W: F(1) = F(nz-1), F(nz) = F(2) */
    double *b2 = malloc(plane * sizeof(double)), *be = malloc(plane * sizeof(double));
    memcpy(b2, F + IX(1, 1, 2, n1, n2), plane * sizeof(double));
    memcpy(be, F + IX(1, 1, n3 - 1, n1, n2), plane * sizeof(double));
    memcpy(F + IX(1, 1, 1, n1, n2), be, plane * sizeof(double));
    memcpy(F + IX(1, 1, n3, n1, n2), b2, plane * sizeof(double));
    free(b2); free(be);
  } else { /* centres: F(1)=F(nzg-2), F(nzg-1)=F(2), F(nzg)=F(3) */
    double *b2 = malloc(plane * sizeof(double)), *b3 = malloc(plane * sizeof(double));
    double *be = malloc(plane * sizeof(double));
    memcpy(b2, F + IX(1, 1, 2, n1, n2), plane * sizeof(double));
    memcpy(b3, F + IX(1, 1, 3, n1, n2), plane * sizeof(double));
    memcpy(be, F + IX(1, 1, n3 - 2, n1, n2), plane * sizeof(double));
    memcpy(F + IX(1, 1, 1, n1, n2), be, plane * sizeof(double));
    memcpy(F + IX(1, 1, n3 - 1, n1, n2), b2, plane * sizeof(double));
    memcpy(F + IX(1, 1, n3, n1, n2), b3, plane * sizeof(double));
    free(b2); free(b3); free(be);
  }
}
/* apply_Dirichlet_bc_y (:237-252) */
static void dirichlet_y(double *F, int n1, int n2, int n3, int faces) {
  for (int k = 1; k <= n3; k++)
    for (int i = 1; i <= n1; i++) {
      if (faces) {
        F[IX(i, 1, k, n1, n2)] = 0.0;
        F[IX(i, n2, k, n1, n2)] = 0.0;
      } else {
        F[IX(i, 1, k, n1, n2)] = -F[IX(i, 2, k, n1, n2)];
        F[IX(i, n2, k, n1, n2)] = -F[IX(i, n2 - 1, k, n1, n2)];
      }
    }
}

void orc_apply_bc(const orc_grid *g, double *U, double *V, double *W) {
  int nx = g->nx, ny = g->ny, nz = g->nz, nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
  /* order of src/boundary_conditions.f90:26-48 */
  periodic_x(U, nx, nyg, nzg, 1);
  periodic_x(V, nxg, ny, nzg, 0);
  periodic_x(W, nxg, nyg, nz, 0);
  periodic_z(U, nx, nyg, nzg, 0);
  periodic_z(V, nxg, ny, nzg, 0);
  periodic_z(W, nxg, nyg, nz, 1);
  dirichlet_y(U, nx, nyg, nzg, 0);
  dirichlet_y(V, nxg, ny, nzg, 1);
  dirichlet_y(W, nxg, nyg, nz, 0);
}

/* ------------------------------------------------------------------------ */
/* divergence / gradient: src/differential_operators.f90                    */
/* ------------------------------------------------------------------------ */
void orc_divergence(const orc_grid *g, double *div, const double *U,
                    const double *V, const double *W) {
  int nx = g->nx, ny = g->ny, nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
  const double *x = g->x, *y = g->y, *z = g->z;
  size_t n = (size_t)(nxg - 2) * (nyg - 2) * (nzg - 1);
  memset(div, 0, n * sizeof(double)); /* div_ = 0d0 (:62) */
  for (int k = 2; k <= nzg - 1; k++)
    for (int j = 2; j <= nyg - 1; j++)
      for (int i = 2; i <= nxg - 1; i++) { /* :64-72 */
        double t1 = (U[IX(i, j, k, nx, nyg)] - U[IX(i - 1, j, k, nx, nyg)]) / (x[i - 1] - x[i - 2]);
        double t2 = (V[IX(i, j, k, nxg, ny)] - V[IX(i, j - 1, k, nxg, ny)]) / (y[j - 1] - y[j - 2]);
        double t3 = (W[IX(i, j, k, nxg, nyg)] - W[IX(i, j, k - 1, nxg, nyg)]) / (z[k - 1] - z[k - 2]);
        div[RX(i, j, k, g)] = t1 + t2 + t3;
      }
}

void orc_gradient(const orc_grid *g, double *Gx, double *Gy, double *Gz,
                  const double *p) {
  int nx = g->nx, ny = g->ny, nz = g->nz, nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
  const double *xg = g->xg, *yg = g->yg, *zg = g->zg;
  memset(Gx, 0, (size_t)nx * nyg * nzg * sizeof(double)); /* :24-26 */
  memset(Gy, 0, (size_t)nxg * ny * nzg * sizeof(double));
  memset(Gz, 0, (size_t)nxg * nyg * nz * sizeof(double));
  for (int k = 2; k <= nzg - 1; k++) /* :28-34 */
    for (int j = 2; j <= nyg - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        Gx[IX(i, j, k, nx, nyg)] = (p[RX(i + 1, j, k, g)] - p[RX(i, j, k, g)]) / (xg[i] - xg[i - 1]);
  for (int k = 2; k <= nzg - 1; k++) /* :36-42 */
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        Gy[IX(i, j, k, nxg, ny)] = (p[RX(i, j + 1, k, g)] - p[RX(i, j, k, g)]) / (yg[j] - yg[j - 1]);
  for (int k = 2; k <= nz - 1; k++) /* :44-50 */
    for (int j = 2; j <= nyg - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        Gz[IX(i, j, k, nxg, nyg)] = (p[RX(i, j, k + 1, g)] - p[RX(i, j, k, g)]) / (zg[k] - zg[k - 1]);
}

double orc_max_abs_div(const orc_grid *g, const double *U, const double *V,
                       const double *W) {
  size_t n = (size_t)(g->nxg - 2) * (g->nyg - 2) * (g->nzg - 1);
  double *d = malloc(n * sizeof(double));
  orc_divergence(g, d, U, V, W);
  double m = 0.0;
  for (size_t i = 0; i < n; i++) if (fabs(d[i]) > m) m = fabs(d[i]);
  free(d);
  return m;
}

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, unnormalised.  Stands in for the     */
/* FFTW 3.3.10 C2C transforms of src/poisson.f90:45,80 (same mathematical    */
/* DFT; FFTW's plan-dependent round-off is not reproducible).                */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, nrad, rad[32];
  cplx *tw; /* tw[k] = exp(-2 pi i k / n) */
} fft_plan;

static int fft_plan_init(fft_plan *p, int n) {
  p->n = n;
  p->nrad = 0;
  int m = n;
  while (m % 4 == 0 && p->nrad < 32) { p->rad[p->nrad++] = 4; m /= 4; }
  while (m % 2 == 0 && p->nrad < 32) { p->rad[p->nrad++] = 2; m /= 2; }
  for (int f = 3; m > 1 && p->nrad < 32; f += 2)
    while (m % f == 0 && p->nrad < 32) { p->rad[p->nrad++] = f; m /= f; }
  if (m != 1) return -1;
  p->tw = malloc((size_t)n * sizeof(cplx));
  /* twiddles evaluated in extended precision and rounded once */
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; k++) {
    long double a = -2.0L * pi * (long double)k / (long double)n;
    p->tw[k].re = (double)cosl(a);
    p->tw[k].im = (double)sinl(a);
  }
  return 0;
}
static void fft_plan_free(fft_plan *p) { free(p->tw); p->tw = NULL; }

/* small DFT of radix R with direction sign (twiddle table of the plan) */
static void dft_small(cplx *v, int R, int sign, const fft_plan *p) {
  if (R == 2) {
    cplx a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  } else if (R == 4) {
    cplx a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cplx a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    /* a3 = d * (sign * i) */
    cplx a3 = {-sign * d.im, sign * d.re};
    v[0] = cadd(a0, a2); v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3); v[3] = csub(a1, a3);
  } else if (R == 3) {
    const double s3 = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
    cplx t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cplx m = {v[0].re - 0.5 * t.re, v[0].im - 0.5 * t.im};
    cplx r = {-sign * s3 * d.im, sign * s3 * d.re}; /* sign*i*s3*d */
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else { /* generic O(R^2) */
    cplx out[64];
    int n = p->n, step = n / R;
    for (int m = 0; m < R; m++) {
      cplx acc = {0.0, 0.0};
      for (int r = 0; r < R; r++) {
        cplx w = p->tw[(size_t)((long)r * m % R) * step];
        if (sign > 0) w.im = -w.im;
        acc = cadd(acc, cmul(v[r], w));
      }
      out[m] = acc;
    }
    for (int m = 0; m < R; m++) v[m] = out[m];
  }
}

/* x (length n) -> transformed in place; work has length n */
static void fft_exec(const fft_plan *p, int sign, cplx *x, cplx *work) {
  int n = p->n, Ns = 1;
  cplx *src = x, *dst = work;
  for (int s = 0; s < p->nrad; s++) {
    int R = p->rad[s], stride = n / R;
    cplx v[64];
    for (int j = 0; j < n / R; j++) {
      int k = j % Ns;
      for (int r = 0; r < R; r++) {
        cplx a = src[j + r * stride];
        /* w^(r k) with w = exp(sign 2 pi i / (Ns R)) */
        size_t ti = ((size_t)r * k * (size_t)(n / (Ns * R))) % (size_t)n;
        cplx w = p->tw[ti];
        if (sign > 0) w.im = -w.im;
        v[r] = (r == 0 || k == 0) ? a : cmul(a, w);
      }
      dft_small(v, R, sign, p);
      int idx = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[idx + r * Ns] = v[r];
    }
    cplx *t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, (size_t)n * sizeof(cplx));
}

int orc_fft(int n, int sign, double *data) {
  fft_plan p;
  if (n < 1 || fft_plan_init(&p, n)) return -1;
  for (int s = 0; s < p.nrad; s++) if (p.rad[s] > 64) { fft_plan_free(&p); return -1; }
  cplx *w = malloc((size_t)n * sizeof(cplx));
  fft_exec(&p, sign < 0 ? -1 : 1, (cplx *)data, w);
  free(w);
  fft_plan_free(&p);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* zgtsv restated (LAPACK 3.x reference algorithm: Gaussian elimination with */
/* partial pivoting, row interchange when CABS1(D(k)) < CABS1(DL(k))).       */
/* Call site: src/poisson.f90:72.                                            */
/* ------------------------------------------------------------------------ */
static int zgtsv_c(int n, cplx *dl, cplx *d, cplx *du, cplx *b) {
  if (n <= 0) return 0;
  for (int k = 0; k < n - 1; k++) {
    if (dl[k].re == 0.0 && dl[k].im == 0.0) {
      if (d[k].re == 0.0 && d[k].im == 0.0) return k + 1;
    } else if (cabs1(d[k]) >= cabs1(dl[k])) {
      /* no row interchange */
      cplx mult = cdiv(dl[k], d[k]);
      d[k + 1] = csub(d[k + 1], cmul(mult, du[k]));
      b[k + 1] = csub(b[k + 1], cmul(mult, b[k]));
      if (k < n - 2) { dl[k].re = 0.0; dl[k].im = 0.0; }
    } else {
      /* interchange rows k and k+1 */
      cplx mult = cdiv(d[k], dl[k]);
      d[k] = dl[k];
      cplx temp = d[k + 1];
      d[k + 1] = csub(du[k], cmul(mult, temp));
      if (k < n - 2) {
        dl[k] = du[k + 1];
        cplx t2 = cmul(mult, dl[k]);
        du[k + 1].re = -t2.re; du[k + 1].im = -t2.im;
      }
      du[k] = temp;
      temp = b[k];
      b[k] = b[k + 1];
      b[k + 1] = csub(temp, cmul(mult, b[k + 1]));
    }
  }
  if (d[n - 1].re == 0.0 && d[n - 1].im == 0.0) return n;
  /* back solve with U */
  b[n - 1] = cdiv(b[n - 1], d[n - 1]);
  if (n > 1) b[n - 2] = cdiv(csub(b[n - 2], cmul(du[n - 2], b[n - 1])), d[n - 2]);
  for (int k = n - 3; k >= 0; k--)
    b[k] = cdiv(csub(csub(b[k], cmul(du[k], b[k + 1])), cmul(dl[k], b[k + 2])), d[k]);
  return 0;
}

int orc_zgtsv(int n, double *dl, double *d, double *du, double *b) {
  return zgtsv_c(n, (cplx *)dl, (cplx *)d, (cplx *)du, (cplx *)b);
}

/* Thomas algorithm without pivoting on the same real-coefficient system
 * (diagnostic variant; rounding differs from zgtsv only where it pivots). */
static void thomas_real(int n, const double *dl, const double *d,
                        const double *du, cplx *b, double *cp) {
  double den = d[0];
  cp[0] = (n > 1) ? du[0] / den : 0.0;
  b[0].re /= den; b[0].im /= den;
  for (int k = 1; k < n; k++) {
    den = d[k] - dl[k - 1] * cp[k - 1];
    if (k < n - 1) cp[k] = du[k] / den;
    b[k].re = (b[k].re - dl[k - 1] * b[k - 1].re) / den;
    b[k].im = (b[k].im - dl[k - 1] * b[k - 1].im) / den;
  }
  for (int k = n - 2; k >= 0; k--) {
    b[k].re -= cp[k] * b[k + 1].re;
    b[k].im -= cp[k] * b[k + 1].im;
  }
}

/* ------------------------------------------------------------------------ */
/* solve_poisson_equation: src/poisson.f90:33-101                            */
/* ------------------------------------------------------------------------ */
void orc_poisson(const orc_grid *g, double *rhs, int rk_step, double dt,
                 double *P, int tri_mode, int nthreads) {
  int nxp = g->nxp, nzp = g->nzp, nym = g->nym;
  int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
  size_t nplane = (size_t)nxp * nzp;
  cplx *hat = malloc(nplane * nym * sizeof(cplx)); /* rhs_hat(0:mx, 2:nyg-1, 0:mz) */
  fft_plan px, pz;
  fft_plan_init(&px, nxp);
  fft_plan_init(&pz, nzp);
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif

  /* forward 2-D DFT of every y plane (:43-47) */
#pragma omp parallel num_threads(nt)
  {
    cplx *plane = malloc(nplane * sizeof(cplx));
    int nmax = nxp > nzp ? nxp : nzp;
    cplx *col = malloc(nmax * sizeof(cplx)), *work = malloc(nmax * sizeof(cplx));
#pragma omp for schedule(static)
    for (int j = 2; j <= nyg - 1; j++) {
      for (int k = 2; k <= nzp + 1; k++)
        for (int i = 2; i <= nxp + 1; i++) {
          cplx c = {rhs[RX(i, j, k, g)], 0.0}; /* dcmplx(rhs) */
          plane[(size_t)(k - 2) * nxp + (i - 2)] = c;
        }
      for (int k = 0; k < nzp; k++) fft_exec(&px, -1, plane + (size_t)k * nxp, work);
      for (int i = 0; i < nxp; i++) {
        for (int k = 0; k < nzp; k++) col[k] = plane[(size_t)k * nxp + i];
        fft_exec(&pz, -1, col, work);
        for (int k = 0; k < nzp; k++) hat[((size_t)k * nym + (j - 2)) * nxp + i] = col[k];
      }
    }
    free(plane); free(col); free(work);
  }

  /* tridiagonal solve per (kx,kz) mode (:50-75) */
#pragma omp parallel num_threads(nt)
  {
    cplx *D = malloc(nym * sizeof(cplx)), *DL = malloc(nym * sizeof(cplx));
    cplx *DU = malloc(nym * sizeof(cplx)), *B = malloc(nym * sizeof(cplx));
    double *dr = malloc(nym * sizeof(double)), *cp = malloc(nym * sizeof(double));
#pragma omp for schedule(static)
    for (int k = 0; k < nzp; k++)
      for (int i = 0; i < nxp; i++) {
        for (int jj = 0; jj < nym; jj++) {
          /* D(j) = Dyy(j,j) + kxx(i) + kzz(k) */
          D[jj].re = (g->dyy_d[jj] + g->kxx[i]) + g->kzz[k];
          D[jj].im = 0.0;
          dr[jj] = D[jj].re;
        }
        for (int jj = 0; jj < nym - 1; jj++) {
          DL[jj].re = g->dyy_l[jj]; DL[jj].im = 0.0;
          DU[jj].re = g->dyy_u[jj]; DU[jj].im = 0.0;
        }
        if (i == 0 && k == 0) { D[0].re = 3.0 / 2.0 * D[0].re; dr[0] = D[0].re; } /* :69 */
        for (int jj = 0; jj < nym; jj++) B[jj] = hat[((size_t)k * nym + jj) * nxp + i];
        if (tri_mode == 0)
          (void)zgtsv_c(nym, DL, D, DU, B); /* INFO ignored, as at :72 */
        else
          thomas_real(nym, g->dyy_l, dr, g->dyy_u, B, cp);
        for (int jj = 0; jj < nym; jj++) hat[((size_t)k * nym + jj) * nxp + i] = B[jj];
      }
    free(D); free(DL); free(DU); free(B); free(dr); free(cp);
  }

  /* inverse 2-D DFT, normalise, real part (:78-82) */
  const double norm = (double)((long)nxp * (long)nzp);
#pragma omp parallel num_threads(nt)
  {
    cplx *plane = malloc(nplane * sizeof(cplx));
    int nmax = nxp > nzp ? nxp : nzp;
    cplx *col = malloc(nmax * sizeof(cplx)), *work = malloc(nmax * sizeof(cplx));
#pragma omp for schedule(static)
    for (int j = 2; j <= nyg - 1; j++) {
      for (int i = 0; i < nxp; i++) {
        for (int k = 0; k < nzp; k++) col[k] = hat[((size_t)k * nym + (j - 2)) * nxp + i];
        fft_exec(&pz, +1, col, work);
        for (int k = 0; k < nzp; k++) plane[(size_t)k * nxp + i] = col[k];
      }
      for (int k = 0; k < nzp; k++) fft_exec(&px, +1, plane + (size_t)k * nxp, work);
      for (int k = 2; k <= nzp + 1; k++)
        for (int i = 2; i <= nxp + 1; i++)
          rhs[RX(i, j, k, g)] = plane[(size_t)(k - 2) * nxp + (i - 2)].re / norm;
    }
    free(plane); free(col); free(work);
  }
  free(hat);
  fft_plan_free(&px);
  fft_plan_free(&pz);

  /* apply_periodic_xz_centers (:106-125), one domain: x then z */
  for (int k = 2; k <= nzg; k++)
    for (int j = 2; j <= nyg - 1; j++) rhs[RX(nxg - 1, j, k, g)] = rhs[RX(2, j, k, g)];
  for (int j = 2; j <= nyg - 1; j++)
    for (int i = 2; i <= nxg - 1; i++) rhs[RX(i, j, nzp + 2, g)] = rhs[RX(i, j, 2, g)];
  /* update_ghost_interior_planes_centers (:130-156): no neighbour on one
   * domain, plane nzg keeps its value. */

  /* pressure side effect (:91-99).  The section P(2:nxg-1,2:nyg-1,2:nzg-1)
   * takes the first nzg-2 z planes of rhs. */
  if (rk_step == 1 && P) {
    const double c = dt * (8.0 / 15.0); /* dt*rk_coef(1,1) */
    int nx1 = nxg, ny1 = nyg;
    for (int k = 2; k <= nzg - 1; k++)
      for (int j = 2; j <= nyg - 1; j++)
        for (int i = 2; i <= nxg - 1; i++) P[IX(i, j, k, nx1, ny1)] = rhs[RX(i, j, k, g)] / c;
    for (int k = 1; k <= nzg; k++)
      for (int j = 1; j <= nyg; j++) {
        P[IX(1, j, k, nx1, ny1)] = P[IX(2, j, k, nx1, ny1)];
        P[IX(nxg, j, k, nx1, ny1)] = P[IX(nxg - 1, j, k, nx1, ny1)];
      }
    for (int k = 1; k <= nzg; k++)
      for (int i = 1; i <= nxg; i++) {
        P[IX(i, 1, k, nx1, ny1)] = P[IX(i, 2, k, nx1, ny1)];
        P[IX(i, nyg, k, nx1, ny1)] = P[IX(i, nyg - 1, k, nx1, ny1)];
      }
    size_t pl = (size_t)nxg * nyg;
    memcpy(P + IX(1, 1, 1, nx1, ny1), P + IX(1, 1, 2, nx1, ny1), pl * sizeof(double));
    memcpy(P + IX(1, 1, nzg, nx1, ny1), P + IX(1, 1, nzg - 1, nx1, ny1), pl * sizeof(double));
  }
}

/* ------------------------------------------------------------------------ */
/* compute_non_IB_projection: src/projection.f90:22-46                       */
/* ------------------------------------------------------------------------ */
void orc_projection(const orc_grid *g, double *U, double *V, double *W,
                    double *Uint, double *Vint, double *Wint, double *Pint,
                    double *rhs, double *P, int rk_step, double dt,
                    int tri_mode, int nthreads) {
  size_t nu = (size_t)g->nx * g->nyg * g->nzg;
  size_t nv = (size_t)g->nxg * g->ny * g->nzg;
  size_t nw = (size_t)g->nxg * g->nyg * g->nz;
  size_t nr = (size_t)(g->nxg - 2) * (g->nyg - 2) * (g->nzg - 1);
  double *ui = Uint ? Uint : malloc(nu * sizeof(double));
  double *vi = Vint ? Vint : malloc(nv * sizeof(double));
  double *wi = Wint ? Wint : malloc(nw * sizeof(double));
  memcpy(ui, U, nu * sizeof(double)); /* :26-28 */
  memcpy(vi, V, nv * sizeof(double));
  memcpy(wi, W, nw * sizeof(double));
  orc_divergence(g, rhs, U, V, W);                        /* :30 */
  orc_poisson(g, rhs, rk_step, dt, P, tri_mode, nthreads); /* :32 */
  for (size_t q = 0; q < nr; q++) rhs[q] = rhs[q] / dt;   /* :33 */
  if (Pint) memcpy(Pint, rhs, nr * sizeof(double));      /* :36 */
  for (size_t q = 0; q < nr; q++) rhs[q] = dt * rhs[q];   /* :38 */
  orc_gradient(g, U, V, W, rhs);                         /* :39 */
  for (size_t q = 0; q < nu; q++) U[q] = ui[q] - U[q];  /* :42-44 */
  for (size_t q = 0; q < nv; q++) V[q] = vi[q] - V[q];
  for (size_t q = 0; q < nw; q++) W[q] = wi[q] - W[q];
  if (!Uint) free(ui);
  if (!Vint) free(vi);
  if (!Wint) free(wi);
}
