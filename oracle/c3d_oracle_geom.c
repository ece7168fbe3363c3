/* CPU oracle (test infrastructure): marker geometry of the wall bodies of
 * setup_IB_geometry (src/immersed_boundary_geometry.f90:70-390), written
 * loop for loop in the reference's operation order so that the HIP
 * generator can be compared with it to the last bit where the math
 * library's sin/cos agree. */
#include <math.h>
#include <stdlib.h>

#include "c3d_oracle.h"
#include "c3d_oracle_impl.h"

/* pi of src/global.f90:34 */
static double ref_pi(void) { return 4.0 * atan(1.0); }

static double sq(double a) { return a * a; }

/* sb of one wall: dzb * (a1/2 + a2/2), a1 and a2 the chord lengths to the
 * x-periodic neighbours in the same slab and wall (:234-262, :346-378). */
static void wall_sb(int nxb, int first, double Lxp, double dzb, const double *xb,
                    const double *yb, double *sb) {
  for (int i = 1; i <= nxb; i++) {
    const int k = first + i - 1, kl = first + nxb - 1, kf = first;
    double a1, a2;
    if (i == 1) {
      a1 = sqrt(sq(xb[k] - (xb[kl] - Lxp)) + sq(yb[k] - yb[kl]));
      a2 = sqrt(sq(xb[k] - xb[k + 1]) + sq(yb[k] - yb[k + 1]));
    } else if (i == nxb) {
      a1 = sqrt(sq(xb[k] - xb[k - 1]) + sq(yb[k] - yb[k - 1]));
      a2 = sqrt(sq(xb[k] - (xb[kf] + Lxp)) + sq(yb[k] - yb[kf]));
    } else {
      a1 = sqrt(sq(xb[k] - xb[k - 1]) + sq(yb[k] - yb[k - 1]));
      a2 = sqrt(sq(xb[k] - xb[k + 1]) + sq(yb[k] - yb[k + 1]));
    }
    sb[k] = dzb * (0.5 * a1 + 0.5 * a2);
  }
}

int orc_wave_geometry(const orc_grid *g, int body, int nxb, int nzb, double p1, double p2,
                      double p3, double ramp, double t, double Lyc, double *xb, double *yb,
                      double *zb, double *normals, double *sb, double *ub, int *yref) {
  if (body != 1 && body != 2) return -1;
  if (nxb < 3 || nzb < 1) return -1;
  const double pi = ref_pi(), Lxp = g->Lxp, Lzp = g->Lzp;
  const double dxb = Lxp / (double)nxb, dzb = Lzp / (double)nzb; /* compute_nb :41-52 */
  const int nb = 2 * nxb * nzb, ny = g->ny;
  for (int q = 0; q < 3 * nb; q++) normals[q] = 0.0, ub[q] = 0.0;
  double amp = p1, damp_dt = 0.0;
  if (body == 2 && t < ramp && ramp > 0.0) { /* :274-281 */
    const double tau = t / ramp;
    amp = p1 * 0.5 * (1.0 - cos(pi * tau));
    damp_dt = p1 * 0.5 * (pi / ramp) * sin(pi * tau);
  }
  const double phi = pi;
  for (int j = 1; j <= nzb; j++)
    for (int i = 1; i <= nxb; i++) {
      const int k = (i + 2 * nxb * (j - 1)) - 1, kt = k + nxb;
      xb[k] = ((double)i - 0.5) * dxb;
      xb[kt] = ((double)i - 0.5) * dxb;
      zb[k] = ((double)j - 0.5) * dzb;
      zb[kt] = ((double)j - 0.5) * dzb;
      yref[k] = 1;
      yref[kt] = ny;
      double s_b, s_t;
      if (body == 1) { /* :186-191, :219-232 */
        yb[k] = p1 * sin(2.0 * pi * p3 * xb[k] / Lxp) * cos(p2 * t);
        yb[kt] = Lyc + p1 * sin(2.0 * pi * p3 * xb[k] / Lxp) * cos(p2 * t);
        const double phase = 2.0 * pi * p3 * xb[k] / Lxp;
        ub[nb + k] = -(p1 * p2 * sin(phase) * sin(p2 * t));
        ub[nb + kt] = -(p1 * p2 * sin(phase) * sin(p2 * t));
        s_b = p1 * 2.0 * pi * p3 / Lxp * cos(phase) * cos(p2 * t);
        s_t = p1 * 2.0 * pi * p3 / Lxp * cos(phase) * cos(p2 * t);
      } else { /* :290-337 */
        const double phase = p3 * (xb[k] - p2 * t);
        yb[k] = amp / (p2 * p3) * sin(phase);
        yb[kt] = Lyc + amp / (p2 * p3) * sin(phase + phi);
        ub[nb + k] = -amp * cos(phase) + (damp_dt / (p2 * p3)) * sin(phase);
        ub[nb + kt] = -amp * cos(phase + phi) + (damp_dt / (p2 * p3)) * sin(phase + phi);
        s_b = amp / p2 * cos(phase);
        s_t = amp / p2 * cos(phase + phi);
      }
      /* tangents_1 (orthogonal to z) and the normals built from them */
      const double t1b = 1.0 / sqrt(1.0 + s_b * s_b), t1t = 1.0 / sqrt(1.0 + s_t * s_t);
      const double t2b = s_b / sqrt(1.0 + s_b * s_b), t2t = s_t / sqrt(1.0 + s_t * s_t);
      normals[k] = t2b;
      normals[kt] = -t2t;
      normals[nb + k] = -t1b;
      normals[nb + kt] = t1t;
    }
  for (int j = 1; j <= nzb; j++) {
    const int first = 2 * nxb * (j - 1);
    wall_sb(nxb, first, Lxp, dzb, xb, yb, sb);
    if (body == 1) /* :264: the top wall copies the bottom wall's areas */
      for (int i = 0; i < nxb; i++) sb[first + nxb + i] = sb[first + i];
    else
      wall_sb(nxb, first + nxb, Lxp, dzb, xb, yb, sb);
  }
  /* stop condition of :199-215 / :305-321: support inside the grid */
  for (int k = 0; k < nb; k++) {
    const int top = (k % (2 * nxb)) >= nxb;
    if (!top && yb[k] < g->y[2]) return k + 1;
    if (top && yb[k] > g->y[ny - 3]) return k + 1;
  }
  return 0;
}
