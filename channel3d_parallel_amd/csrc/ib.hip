// Immersed-boundary kernels for gfx950: stencil set-up, interpolation,
// spreading, Schur-complement pieces and BiCGSTAB vector steps.
//
// Interpolation runs one wavefront per marker: the 125 stencil points are
// spread over the 64 lanes and reduced with a fixed butterfly, so results are
// deterministic.  Spreading is a gather over a node -> entry map built once
// per static body, summing each node's contributions in ascending marker
// order (the order of the reference's serial scatter), with no atomics.
#include <hip/hip_runtime.h>

#include "ib.hpp"

#pragma clang fp contract(off)

namespace c3d {
namespace ib {

namespace {

constexpr int kBlock = 256;

// deltafnc, src/immersed_boundary_operators.f90:270-294.
__device__ double deltafnc(double r, double dr) {
  const double pi = 4.0 * atan(1.0);
  const double s3 = sqrt(3.0);
  const double r1 = fabs(r) / dr;
  const double r2 = r1 * r1;
  double d;
  if (r1 <= 1.0) {
    d = 17.0 / 48.0 + s3 * pi / 108.0 + r1 / 4.0 - r2 / 4.0 +
        (1.0 - 2.0 * r1) / 16.0 * sqrt(-12.0 * r2 + 12.0 * r1 + 1.0) -
        s3 / 12.0 * asin(s3 / 2.0 * (2.0 * r1 - 1.0));
  } else if (r1 <= 2.0) {
    d = 55.0 / 48.0 - s3 * pi / 108.0 - 13.0 * r1 / 12.0 + r2 / 4.0 +
        (2.0 * r1 - 3.0) / 48.0 * sqrt(-12.0 * r2 + 36.0 * r1 - 23.0) +
        s3 / 36.0 * asin(s3 / 2.0 * (2.0 * r1 - 3.0));
  } else {
    d = 0.0;
  }
  return d / dr;
}

__device__ __forceinline__ int floor_div(int a, int b) {
  return (int)floor((double)a / (double)b);
}

// One thread per (marker, stencil): pivots (src/immersed_boundary_operators.f90:28-42)
// and the per-axis windows of the u (:46-98), v (:101-148), w (:151-194) and
// centre (:197-244) stencils on one periodic domain, centre z index nzg-1
// stored as its canonical plane 2.
__global__ void k_setup(DevGrid g, SetupGrid sg, Markers m, int* bad) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = m.nb;
  if (tid >= 4 * nb) return;
  const int s = tid / nb, l = tid - s * nb;
  const int nx = g.nx, nz = g.nz, nzg = g.nzg;
  const double dx = sg.dx, dz = sg.dz, dymin = sg.dymin;
  const double xb = m.xb[l], yb = m.yb[l], zb = m.zb[l];
  const double* x = sg.x;
  const double* y = sg.y;
  const double* z = sg.z;
  const int x_piv = (int)floor(xb / dx) + 1;
  const int xm_piv = (int)floor((xb + 0.5 * dx) / dx);
  const int yr = m.yref ? m.yref[l] : 1;
  const int y_piv = (int)floor((yb - y[yr - 1]) / dymin) + yr;
  const int ym_piv = (int)floor((yb - (y[yr - 1] + 0.5 * dymin)) / dymin) + yr;
  const int ylim = (s == 1) ? g.ny - 1 : g.nyg - 1;
  const int z_piv = (int)floor(zb / dz) + 1;
  const int zm_piv = (int)floor((zb + 0.5 * dz) / dz) + 1;
  Stencil st = m.st[s];
  for (int t = 0; t < 5; t++) {
    const int o = t - 2;
    // x
    int X;
    double xgd;
    if (s == 0) {
      const int ii = x_piv + o;
      const int xs = floor_div(ii - 1, nx - 2);
      const int iip = ii - xs * (nx - 2);
      X = (iip == 1) ? nx - 1 : iip;
      xgd = x[iip - 1] + xs * sg.Lxp;
    } else {
      const int ii = xm_piv + o - (s == 2 ? 1 : 0);
      const int xs = floor_div(ii - 1, nx - 2);
      const int iip = ii - xs * (nx - 2);
      X = iip + 1;
      xgd = 0.5 * (x[iip - 1] + x[iip]) + xs * sg.Lxp;
    }
    // y (a window reaching past the walls is flagged with Y = -1 and
    // rejected by the host before any other IB kernel runs)
    int Y = -1;
    double ygd = 0.0;
    if (s == 1) {
      const int jj = y_piv + o;
      if (jj >= 1 && jj <= g.ny) {
        Y = jj;
        ygd = y[jj - 1];
      }
    } else {
      const int jj = ym_piv + o;
      if (jj >= 1 && jj <= g.ny - 1) {
        Y = jj + 1;
        ygd = 0.5 * (y[jj - 1] + y[jj]);
      }
    }
    // z
    int Z;
    double zgd;
    if (s == 2) {
      const int kk = z_piv + o;
      const int zs = floor_div(kk - 2, nz - 2);
      const int kkp = kk - zs * (nz - 2);
      Z = kkp;
      zgd = z[kkp - 1] + zs * sg.Lzp;
    } else {
      const int kk = zm_piv + o;
      const int zs = floor_div(kk - 2, nz - 1);
      const int kkp = kk - zs * (nz - 2);
      Z = (kkp == nzg - 1) ? 2 : kkp;
      zgd = 0.5 * (z[kkp - 2] + z[kkp - 1]) + zs * sg.Lzp;
    }
    const double ddx = xgd - xb, ddy = ygd - yb, ddz = zgd - zb;
    if (bad && (X < 2 || X > nx - 1 || Z < 2 || Z > nz - 1 || Y < 2 || Y > ylim))
      atomicMax(bad, l + 1);
    st.idx[(0 * 5 + t) * nb + l] = X;
    st.idx[(1 * 5 + t) * nb + l] = Y;
    st.idx[(2 * 5 + t) * nb + l] = Z;
    st.dd[(0 * 5 + t) * nb + l] = ddx;
    st.dd[(1 * 5 + t) * nb + l] = ddy;
    st.dd[(2 * 5 + t) * nb + l] = ddz;
    st.phi[(0 * 5 + t) * nb + l] = deltafnc(ddx, dx);
    st.phi[(1 * 5 + t) * nb + l] = deltafnc(ddy, dymin);
    st.phi[(2 * 5 + t) * nb + l] = deltafnc(ddz, dz);
  }
}

// Wall height of marker i (1-based, x-periodic) of one wall at the time wt
// (src/immersed_boundary_geometry.f90:186-191 standing, :295-299 traveling),
// and x itself.  The standing wave's top wall returns the bottom wall's
// curve: its areas are copied from the bottom wall (:264).
__device__ __forceinline__ void wall_point(const WaveBody& w, const WaveTime& wt, int i, bool top,
                                           double& x, double& y) {
  const double dxb = w.Lxp / (double)w.nxb;
  x = ((double)i - 0.5) * dxb;
  if (w.body == 1) {
    y = w.p1 * sin(2.0 * w.pi * w.p3 * x / w.Lxp) * wt.ct;
  } else {
    const double phase = w.p3 * (x - w.p2 * wt.t);
    y = wt.amp / (w.p2 * w.p3) * sin(top ? phase + w.pi : phase);
  }
}

// One thread per marker: position, velocity, normal, area and y_ref_index of
// the moving walls; the area sb = dzb (a1/2 + a2/2) from the chords to the
// x-periodic neighbours on the same wall (:234-262, :346-378).
__global__ void k_wave_geometry(WaveBody w, WaveTime wt, double* __restrict__ xb,
                                double* __restrict__ yb, double* __restrict__ zb,
                                double* __restrict__ normals, double* __restrict__ sb,
                                double* __restrict__ ub, int* __restrict__ yref, int* bad) {
  const int nxb = w.nxb, nb = 2 * nxb * w.nzb;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int j = k / (2 * nxb) + 1, r = k - (j - 1) * 2 * nxb;
  const bool top = r >= nxb;
  const int i = (top ? r - nxb : r) + 1;
  const double dzb = w.Lzp / (double)w.nzb;
  // own point, and the curve the neighbours' chords are measured on
  const bool curve_top = top && w.body == 2;
  double x, yc;
  wall_point(w, wt, i, curve_top, x, yc);
  const double y = top ? w.Lyc + yc : yc;
  double s, vy;
  if (w.body == 1) {
    const double phase = 2.0 * w.pi * w.p3 * x / w.Lxp;
    vy = -(w.p1 * w.p2 * sin(phase) * wt.st);
    s = w.p1 * 2.0 * w.pi * w.p3 / w.Lxp * cos(phase) * wt.ct;
  } else {
    const double phase = w.p3 * (x - w.p2 * wt.t);
    const double ph = top ? phase + w.pi : phase;
    vy = -wt.amp * cos(ph) + (wt.damp / (w.p2 * w.p3)) * sin(ph);
    s = wt.amp / w.p2 * cos(ph);
  }
  const double t1 = 1.0 / sqrt(1.0 + s * s), t2 = s / sqrt(1.0 + s * s);
  // chords on the curve (the top wall of a traveling wave on its own curve)
  double xp, yp, xn, yn;
  wall_point(w, wt, i == 1 ? nxb : i - 1, curve_top, xp, yp);
  wall_point(w, wt, i == nxb ? 1 : i + 1, curve_top, xn, yn);
  const double ys = curve_top ? w.Lyc + yc : yc;
  if (curve_top) {
    yp = w.Lyc + yp;
    yn = w.Lyc + yn;
  }
  if (i == 1) xp = xp - w.Lxp;
  if (i == nxb) xn = xn + w.Lxp;
  const double a1 = sqrt((x - xp) * (x - xp) + (ys - yp) * (ys - yp));
  const double a2 = sqrt((x - xn) * (x - xn) + (ys - yn) * (ys - yn));
  xb[k] = x;
  yb[k] = y;
  zb[k] = ((double)j - 0.5) * dzb;
  sb[k] = dzb * (0.5 * a1 + 0.5 * a2);
  yref[k] = top ? w.ny : 1;
  normals[k] = top ? -t2 : t2;
  normals[nb + k] = top ? t1 : -t1;
  normals[2 * nb + k] = 0.0;
  ub[k] = 0.0;
  ub[nb + k] = vy;
  ub[2 * nb + k] = 0.0;
  if (top ? (y > w.yhi) : (y < w.ylo)) atomicMax(bad, k + 1);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Point p = i + 5*(j + 5*k) of marker l in stencil st: node indices, weight
// and (optionally) the normal distance n . (x_node - xb).
struct Pt {
  int X, Y, Z;
  double w, dxn;
};

template <bool DXN>
__device__ __forceinline__ Pt stencil_point(const Markers& m, const Stencil& st, int l, int p) {
  const int nb = m.nb;
  const int i = p % 5, j = (p / 5) % 5, k = p / 25;
  Pt r;
  r.X = st.idx[(0 * 5 + i) * nb + l];
  r.Y = st.idx[(1 * 5 + j) * nb + l];
  r.Z = st.idx[(2 * 5 + k) * nb + l];
  r.w = ((m.A * st.phi[(0 * 5 + i) * nb + l]) * st.phi[(1 * 5 + j) * nb + l]) *
        st.phi[(2 * 5 + k) * nb + l];
  r.dxn = 0.0;
  if (DXN)
    r.dxn = (m.normals[l] * st.dd[(0 * 5 + i) * nb + l] +
             m.normals[nb + l] * st.dd[(1 * 5 + j) * nb + l]) +
            m.normals[2 * nb + l] * st.dd[(2 * 5 + k) * nb + l];
  return r;
}

// Marker handled by this wavefront (XCD-aware: neighbouring markers, whose
// stencils share nodes, run on the same XCD and L2).
__device__ __forceinline__ int wave_marker(int nb) {
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);
  return (int)bx * (kBlock / 64) + (int)(threadIdx.x >> 6);
}

// Sums over the u, v and w stencils of marker l at once: the 375 points are
// spread over the 64 lanes (up to 6 per lane, all loads in flight together)
// and term(c, point) is accumulated per component; the three sums are then
// reduced across the wavefront.
// Per-axis stencil values of marker l held one per lane: lane 15c + 5a + t
// (c < ncomp) holds entry (axis a, position t) of stencil st[c0 + c]; the
// 125 points then fetch theirs by cross-lane reads instead of 6 loads each.
struct LaneStencil {
  int idx;
  double phi, dd;
};

template <bool DXN>
__device__ __forceinline__ LaneStencil load_lane_stencil(const Markers& m, int l, int lane, int c0,
                                                         int ncomp) {
  LaneStencil r{0, 0.0, 0.0};
  if (lane < 15 * ncomp) {
    const int c = c0 + lane / 15, e = lane % 15;  // e = 5a + t
    const Stencil st = c == 0 ? m.st[0] : (c == 1 ? m.st[1] : (c == 2 ? m.st[2] : m.st[3]));
    const size_t o = (size_t)e * m.nb + l;
    r.idx = st.idx[o];
    r.phi = st.phi[o];
    if (DXN) r.dd = st.dd[o];
  }
  return r;
}

// Point p = i + 5*(j + 5*k) of the c-th loaded stencil, from the lanes.
template <bool DXN>
__device__ __forceinline__ Pt lane_point(const Markers& m, const LaneStencil& ls, int l, int c, int p) {
  const int i = p % 5, j = (p / 5) % 5, k = p / 25;
  const int b = 15 * c;
  Pt r;
  r.X = __shfl(ls.idx, b + i, 64);
  r.Y = __shfl(ls.idx, b + 5 + j, 64);
  r.Z = __shfl(ls.idx, b + 10 + k, 64);
  r.w = ((m.A * __shfl(ls.phi, b + i, 64)) * __shfl(ls.phi, b + 5 + j, 64)) *
        __shfl(ls.phi, b + 10 + k, 64);
  r.dxn = 0.0;
  if (DXN) {
    r.dxn = (m.normals[l] * __shfl(ls.dd, b + i, 64) + m.normals[m.nb + l] * __shfl(ls.dd, b + 5 + j, 64)) +
            m.normals[2 * m.nb + l] * __shfl(ls.dd, b + 10 + k, 64);
  }
  return r;
}

template <bool DXN, typename Term>
__device__ __forceinline__ void gather3(const Markers& m, int l, int lane, Term term, double out[3]) {
  const LaneStencil ls = load_lane_stencil<DXN>(m, l, lane, 0, 3);
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int p = lane + 64 * q;
    const int c = p >= 250 ? 2 : (p >= 125 ? 1 : 0);
    // all lanes take part in the cross-lane reads; out-of-range points add 0
    const Pt pt = lane_point<DXN>(m, ls, l, c, (p < 375) ? p - 125 * c : 0);
    if (p < 375) {
      const double v = term(c, pt);
      acc[0] += c == 0 ? v : 0.0;
      acc[1] += c == 1 ? v : 0.0;
      acc[2] += c == 2 ? v : 0.0;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) out[c] = wave_sum(acc[c]);
}

// E / E1n of U, V, W (src/immersed_boundary_operators.f90:409-501).
template <bool DXN>
__global__ __launch_bounds__(kBlock) void k_interp_uvw(DevGrid g, Markers m, const double* __restrict__ U,
                                                       const double* __restrict__ V,
                                                       const double* __restrict__ W,
                                                       double* __restrict__ out) {
  const int l = wave_marker(m.nb);
  const int lane = threadIdx.x & 63;
  if (l >= m.nb) return;
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;
  double r[3];
  gather3<DXN>(m, l, lane, [&](int c, const Pt& q) {
    const double f = c == 0 ? U[fidx(q.X, q.Y, q.Z, nx, nyg)]
                            : (c == 1 ? V[fidx(q.X, q.Y, q.Z, nxg, ny)] : W[fidx(q.X, q.Y, q.Z, nxg, nyg)]);
    double v = q.w * f;
    if (DXN) v = v * q.dxn;
    return v;
  }, r);
  if (lane == 0)
    for (int c = 0; c < 3; c++) out[c * m.nb + l] = r[c];
}

// E1n of a centre field in rhs_p layout (:503-526), marker l on one wave.
__device__ __forceinline__ void interp_c_marker(const DevGrid& g, const Markers& m, int l, int lane,
                                                const double* __restrict__ P, double sign,
                                                double* __restrict__ out) {
  const LaneStencil ls = load_lane_stencil<true>(m, l, lane, 3, 1);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int p = lane + 64 * q;
    const Pt pt = lane_point<true>(m, ls, l, 0, p < 125 ? p : 0);
    if (p < 125) acc += (pt.w * P[ridx(pt.X, pt.Y, pt.Z, g.nxg, g.nyg)]) * pt.dxn;
  }
  acc = wave_sum(acc);
  if (lane == 0) out[l] = sign * acc;
}

__global__ __launch_bounds__(kBlock) void k_interp_c(DevGrid g, Markers m, const double* __restrict__ P,
                                                     double sign, double* __restrict__ out) {
  const int l = wave_marker(m.nb);
  if (l >= m.nb) return;
  interp_c_marker(g, m, l, threadIdx.x & 63, P, sign, out);
}

// Schur velocity rows (src/projection.f90:131-142) with the gradient of
// dt*phi evaluated at the stencil nodes; every node lies in the gradient's
// index ranges (checked at set-up), where apply_boundary_conditions does not
// write.
__device__ __forceinline__ void schur_u_marker(const DevGrid& g, const Markers& m, int l, int lane,
                                               const double* __restrict__ phi, double dt,
                                               const double* __restrict__ Ureg,
                                               const double* __restrict__ Vreg,
                                               const double* __restrict__ Wreg,
                                               const double* __restrict__ E1nH,
                                               const double* __restrict__ dudn,
                                               double* __restrict__ Sf) {
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;
  double r[3];
  gather3<false>(m, l, lane, [&](int c, const Pt& q) {
    const double p0 = dt * phi[ridx(q.X, q.Y, q.Z, nxg, nyg)];
    double G, R;
    if (c == 0) {
      G = (dt * phi[ridx(q.X + 1, q.Y, q.Z, nxg, nyg)] - p0) / g.dxg[q.X];
      R = Ureg[fidx(q.X, q.Y, q.Z, nx, nyg)];
    } else if (c == 1) {
      G = (dt * phi[ridx(q.X, q.Y + 1, q.Z, nxg, nyg)] - p0) / g.dyg[q.Y];
      R = Vreg[fidx(q.X, q.Y, q.Z, nxg, ny)];
    } else {
      G = (dt * phi[ridx(q.X, q.Y, q.Z + 1, nxg, nyg)] - p0) / g.dzg[q.Z];
      R = Wreg[fidx(q.X, q.Y, q.Z, nxg, nyg)];
    }
    return q.w * (G - R);
  }, r);
  if (lane == 0)
    for (int c = 0; c < 3; c++)
      Sf[c * m.nb + l] = r[c] - E1nH[c * m.nb + l] * dudn[c * m.nb + l];
}

// Both row groups of the Schur product in one launch, one wave per marker:
// the velocity rows (schur_u_marker), then the pressure row -E1n(phi) into
// Sf[3nb..4nb) (as k_interp_c with sign -1), whose nodes the velocity rows
// have just read.
__global__ __launch_bounds__(kBlock) void k_schur_rows(DevGrid g, Markers m, const double* __restrict__ phi,
                                                       double dt, const double* __restrict__ Ureg,
                                                       const double* __restrict__ Vreg,
                                                       const double* __restrict__ Wreg,
                                                       const double* __restrict__ E1nH,
                                                       const double* __restrict__ dudn,
                                                       double* __restrict__ Sf) {
  const int l = wave_marker(m.nb);
  if (l >= m.nb) return;
  const int lane = threadIdx.x & 63;
  schur_u_marker(g, m, l, lane, phi, dt, Ureg, Vreg, Wreg, E1nH, dudn, Sf);
  interp_c_marker(g, m, l, lane, phi, -1.0, Sf + 3 * (size_t)m.nb);
}

// -E u + ub (src/projection.f90:51-52).
__global__ __launch_bounds__(kBlock) void k_rhs_u(DevGrid g, Markers m, const double* __restrict__ U,
                                                  const double* __restrict__ V,
                                                  const double* __restrict__ W,
                                                  double* __restrict__ out) {
  const int l = wave_marker(m.nb);
  const int lane = threadIdx.x & 63;
  if (l >= m.nb) return;
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;
  double r[3];
  gather3<false>(m, l, lane, [&](int c, const Pt& q) {
    const double f = c == 0 ? U[fidx(q.X, q.Y, q.Z, nx, nyg)]
                            : (c == 1 ? V[fidx(q.X, q.Y, q.Z, nxg, ny)] : W[fidx(q.X, q.Y, q.Z, nxg, nyg)]);
    return q.w * f;
  }, r);
  if (lane == 0)
    for (int c = 0; c < 3; c++) out[c * m.nb + l] = -r[c] + m.ub[c * m.nb + l];
}

__global__ void k_aux(Markers m, const double* __restrict__ f, double dt, double nu,
                      double* __restrict__ aux) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = m.nb;
  if (t >= 3 * nb) return;
  const int l = t % nb;
  aux[t] = dt * nu * f[t] - dt * f[3 * nb + l] * m.normals[t];
}

// Static coefficients of the spreading entries: w*factor*sb[l] per entry.
__global__ void k_spread_coeffs(int nb, double A, double factor, const double* __restrict__ phi,
                                const double* __restrict__ sb, Spread sp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sp.nentries) return;
  const int q = sp.q[e];
  const int l = q / 125, p = q - 125 * l;
  const int i = p % 5, j = (p / 5) % 5, k = p / 25;
  const double w = ((A * phi[(0 * 5 + i) * nb + l]) * phi[(1 * 5 + j) * nb + l]) *
                   phi[(2 * 5 + k) * nb + l];
  sp.ml[e] = l;
  sp.wfs[e] = w * factor * sb[l];
}

// R, src/immersed_boundary_operators.f90:544-603: one thread per support node,
// contributions in ascending marker order.
__device__ __forceinline__ void spread_node(const Spread& sp, int t, const double* __restrict__ f,
                                            double* __restrict__ F) {
  double acc = 0.0;
  const int n = sp.cnt[t];
  // loads of several entries in flight (8 measured 1% faster than 4 per
  // BiCGSTAB iteration); the sum stays in entry order
#pragma unroll 8
  for (int i = 0; i < n; i++) {
    const int e = sp.slot[i] + t;
    acc = acc + sp.wfs[e] * f[sp.ml[e]];
  }
  F[sp.node[t]] = acc;
}

__global__ __launch_bounds__(kBlock) void k_spread(Spread sp, const double* __restrict__ f,
                                                   double* __restrict__ F) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sp.nnodes) return;
  spread_node(sp, t, f, F);
}

// The three components in one launch, block row c = component c, reading
// f_c = f + c*nb.
__global__ __launch_bounds__(kBlock) void k_spread3(Spread s0, Spread s1, Spread s2,
                                                    const double* __restrict__ f, int nb,
                                                    double* __restrict__ F0, double* __restrict__ F1,
                                                    double* __restrict__ F2) {
  const int c = blockIdx.y;
  const Spread sp = c == 0 ? s0 : (c == 1 ? s1 : s2);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sp.nnodes) return;
  spread_node(sp, t, f + (size_t)c * nb, c == 0 ? F0 : (c == 1 ? F1 : F2));
}

// remove_mean_per_body for the two-group slab layout
// (src/immersed_boundary_geometry.f90:400-442), then f -= coef*pj.  Block
// partial sums of the two bodies, then every block re-reduces the partials
// in the same fixed order and applies.
constexpr int kMeanBlocks = 64;

__global__ __launch_bounds__(kBlock) void k_body_sums(Markers m, const double* __restrict__ f,
                                                      double* __restrict__ part) {
  __shared__ double s1[kBlock], s2[kBlock];
  const int nb = m.nb, nxb = m.nxb, nxb1 = m.nxb1;
  double a = 0.0, b = 0.0;
  for (int l = blockIdx.x * kBlock + threadIdx.x; l < nb; l += kMeanBlocks * kBlock) {
    if (l % nxb < nxb1)
      a += f[l];
    else
      b += f[l];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[kMeanBlocks + blockIdx.x] = s2[0];
  }
}

__global__ __launch_bounds__(kBlock) void k_remove_mean(Markers m, double* __restrict__ f,
                                                        const double* __restrict__ part,
                                                        const double* __restrict__ coef,
                                                        const double* __restrict__ pj) {
  const int nb = m.nb, nxb = m.nxb, nxb1 = m.nxb1;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < kMeanBlocks; q++) {
    a += part[q];
    b += part[kMeanBlocks + q];
  }
  const double avg1 = a / (double)(nxb1 * m.nzb);
  const double avg2 = b / (double)((nxb - nxb1) * m.nzb);
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= nb) return;
  double v = f[l] - ((l % nxb < nxb1) ? avg1 : avg2);
  if (coef) v = v - coef[l] * pj[l];
  f[l] = v;
}

// U = U_interim - U_reg - G(dt*(P_interim - phi))  (src/projection.f90:83-92)
__global__ __launch_bounds__(kBlock) void k_ib_correct(DevGrid g, const double* __restrict__ Uint,
                                                       const double* __restrict__ Vint,
                                                       const double* __restrict__ Wint,
                                                       const double* __restrict__ Ureg,
                                                       const double* __restrict__ Vreg,
                                                       const double* __restrict__ Wreg,
                                                       const double* __restrict__ Pint,
                                                       const double* __restrict__ phi, double dt,
                                                       double* __restrict__ U,
                                                       double* __restrict__ V,
                                                       double* __restrict__ W) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 1;
  const int k = blockIdx.z + 1;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  if (i > nxg || j > nyg) return;
  auto q = [&](int a, int b, int c) {
    const size_t o = ridx(a, b, c, nxg, nyg);
    return dt * (Pint[o] - phi[o]);
  };
  const bool core = i >= 2 && i <= nx - 1 && j >= 2 && k >= 2;
  bool bad = false;
  if (i <= nx) {  // U(nx, nyg, nzg)
    const size_t o = fidx(i, j, k, nx, nyg);
    double G = 0.0;
    if (core && j <= nyg - 1 && k <= nzg - 1) G = (q(i + 1, j, k) - q(i, j, k)) / g.dxg[i];
    const double u = Uint[o] - Ureg[o] - G;
    U[o] = u;
    bad |= !isfinite(u);
  }
  if (j <= ny) {  // V(nxg, ny, nzg)
    const size_t o = fidx(i, j, k, nxg, ny);
    double G = 0.0;
    if (core && j <= ny - 1 && k <= nzg - 1) G = (q(i, j + 1, k) - q(i, j, k)) / g.dyg[j];
    const double v = Vint[o] - Vreg[o] - G;
    V[o] = v;
    bad |= !isfinite(v);
  }
  if (k <= nz) {  // W(nxg, nyg, nz)
    const size_t o = fidx(i, j, k, nxg, nyg);
    double G = 0.0;
    if (core && j <= nyg - 1 && k <= nz - 1) G = (q(i, j, k + 1) - q(i, j, k)) / g.dzg[k];
    const double w = Wint[o] - Wreg[o] - G;
    W[o] = w;
    bad |= !isfinite(w);
  }
  if (bad && g.nonfinite) atomicOr(g.nonfinite, 1u);
}

__global__ void k_affine(const double* __restrict__ x, double a, double b, double* __restrict__ y,
                         size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a + b * x[t];
}

// interpolate_x/y/z of src/interpolation.f90 into the heaviside ranges.
__global__ void k_h_interp(DevGrid g, const double* __restrict__ Hc, double* __restrict__ Hu,
                           double* __restrict__ Hv, double* __restrict__ Hw) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 2;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 2;
  const int k = blockIdx.z + 2;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg;
  if (i > nxg - 1 || j > nyg - 1) return;
  const double c = Hc[ridx(i, j, k, nxg, nyg)];
  if (i <= nx - 1) Hu[fidx(i, j, k, nx, nyg)] = 0.5 * (c + Hc[ridx(i + 1, j, k, nxg, nyg)]);
  if (j <= ny - 1) Hv[fidx(i, j, k, nxg, ny)] = 0.5 * (c + Hc[ridx(i, j + 1, k, nxg, nyg)]);
  if (k <= nz) Hw[fidx(i, j, k, nxg, nyg)] = 0.5 * (c + Hc[ridx(i, j, k + 1, nxg, nyg)]);
}

// ---- BiCGSTAB ----
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kBlock / 64; q++) r += sh[q];
  return r;
}

__global__ __launch_bounds__(kBlock) void k_dot(const double* __restrict__ a, const double* __restrict__ b,
                                                const double* __restrict__ a2,
                                                const double* __restrict__ b2, size_t n,
                                                double* __restrict__ part, int slot) {
  __shared__ double sh[2][kBlock / 64];
  double s1 = 0.0, s2 = 0.0;
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)kDotBlocks * kBlock) {
    s1 += a[t] * b[t];
    if (a2) s2 += a2[t] * b2[t];
  }
  const double r1 = block_sum(s1, sh[0]);
  if (threadIdx.x == 0) part[slot * kDotBlocks + blockIdx.x] = r1;
  if (a2) {
    const double r2 = block_sum(s2, sh[1]);
    if (threadIdx.x == 0) part[(slot + 1) * kDotBlocks + blockIdx.x] = r2;
  }
}

__device__ __forceinline__ double sum_part(const double* part, int slot) {
  double r = 0.0;
  for (int b = 0; b < kDotBlocks; b++) r += part[slot * kDotBlocks + b];
  return r;
}

// Per-solve input normalisation of the learned guess: out[0] = 1/rms of the
// n_r entries whose squares are summed in partial slot sr, out[1] = the
// same for slot sf (0 for a zero vector).
__global__ void k_norm_scales(const double* __restrict__ part, int sr, double n_r, int sf, double n_f,
                              double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double qr = sum_part(part, sr), qf = sum_part(part, sf);
  out[0] = qr > 0.0 ? 1.0 / sqrt(qr / n_r) : 0.0;
  out[1] = qf > 0.0 ? 1.0 / sqrt(qf / n_f) : 0.0;
}

// Scalar slots: 0 rho_o, 1 alpha, 2 om, 3 rho_n, 4 bta, 5 err, 6 line-search
// step.  The vector kernels below compute their scalar from the dot partials
// themselves (thread 0 of every block, the same fixed-order sum) and block 0
// stores it; no block of a kernel reads a slot that the same kernel writes.
__global__ void k_bicg_scalar(int step, double* sc, const double* __restrict__ part,
                              double* host_err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  switch (step) {
    case 0:
      sc[0] = 1.0;
      sc[1] = 1.0;
      sc[2] = 1.0;
      break;
    case 4: {
      const double e = sum_part(part, 4);
      sc[5] = e;
      if (host_err) {  // pinned host word read after the stream synchronises
        *host_err = e;
        __threadfence_system();
      }
      break;
    }
    case 5: {  // minimal-residual step along d: a = <r, S d> / <S d, S d>
      const double den = sum_part(part, 3);
      sc[6] = den > 0.0 ? sum_part(part, 2) / den : 0.0;
      break;
    }
  }
}

// rho_n = rhat.r (partial slot 0); bta = (rho_n/rho_o)*(alpha/om);
// p = r + bta*(p - om*v).  Writes sc[3] = rho_n, sc[4] = bta.
// Breakdown guards here and in k_bicg_hs / k_bicg_xr: an exactly zero
// denominator gives a zero step instead of the reference's 0/0
// (src/projection.f90:165-172); a zero right-hand side then converges to
// x = 0 in one iteration.  Nonzero denominators are untouched.
__global__ void k_bicg_p(const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ sc, const double* __restrict__ part,
                         double* __restrict__ p, size_t n) {
  __shared__ double bsh;
  if (threadIdx.x == 0) {
    const double rho_n = sum_part(part, 0);
    const double bta = (sc[0] != 0.0 && sc[2] != 0.0) ? (rho_n / sc[0]) * (sc[1] / sc[2]) : 0.0;
    bsh = bta;
    if (blockIdx.x == 0) {
      sc[3] = rho_n;
      sc[4] = bta;
    }
  }
  __syncthreads();
  const double bta = bsh, om = sc[2];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  p[t] = r[t] + bta * (p[t] - om * v[t]);
}

// alpha = rho_n / rhat.v (partial slot 1); h = x + alpha*p; s = r - alpha*v.
// Writes sc[1] = alpha and sc[0] = rho_n (rho_o of the next iteration).
__global__ void k_bicg_hs(const double* __restrict__ x, const double* __restrict__ p,
                          const double* __restrict__ r, const double* __restrict__ v,
                          double* __restrict__ sc, const double* __restrict__ part,
                          double* __restrict__ h, double* __restrict__ sv, size_t n) {
  __shared__ double ash;
  if (threadIdx.x == 0) {
    const double rho_n = sc[3];
    const double den = sum_part(part, 1);
    const double alpha = den != 0.0 ? rho_n / den : 0.0;
    ash = alpha;
    if (blockIdx.x == 0) {
      sc[1] = alpha;
      sc[0] = rho_n;
    }
  }
  __syncthreads();
  const double alpha = ash;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  h[t] = x[t] + alpha * p[t];
  sv[t] = r[t] - alpha * v[t];
}

// om = t.s / t.t (partial slots 2, 3); x = h + om*s; r = s - om*t; partials
// of r.r into slot 4 and of rhat.r into slot 0 (the next iteration's rho_n).
// Writes sc[2] = om.
__global__ __launch_bounds__(kBlock) void k_bicg_xr(const double* __restrict__ h,
                                                    const double* __restrict__ sv,
                                                    const double* __restrict__ tv,
                                                    const double* __restrict__ rhat,
                                                    double* __restrict__ sc,
                                                    double* __restrict__ part,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    size_t n) {
  __shared__ double sh[2][kBlock / 64];
  __shared__ double osh;
  if (threadIdx.x == 0) {
    const double tt = sum_part(part, 3);
    const double om = tt != 0.0 ? sum_part(part, 2) / tt : 0.0;
    osh = om;
    if (blockIdx.x == 0) sc[2] = om;
  }
  __syncthreads();
  const double om = osh;
  double s4 = 0.0, s0 = 0.0;
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)kDotBlocks * kBlock) {
    x[t] = h[t] + om * sv[t];
    const double rr = sv[t] - om * tv[t];
    r[t] = rr;
    s4 += rr * rr;
    s0 += rhat[t] * rr;
  }
  const double t4 = block_sum(s4, sh[0]);
  const double t0 = block_sum(s0, sh[1]);
  if (threadIdx.x == 0) {
    part[4 * kDotBlocks + blockIdx.x] = t4;
    part[0 * kDotBlocks + blockIdx.x] = t0;
  }
}

__global__ void k_bicg_init(const double* __restrict__ b, double* __restrict__ r,
                            double* __restrict__ rhat, double* __restrict__ p,
                            double* __restrict__ v, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double rr = b ? b[t] - r[t] : r[t];  // b == null: r already holds the residual
  r[t] = rr;
  rhat[t] = rr;
  p[t] = 0.0;
  v[t] = 0.0;
}

// x = x + a d;  r = r - a S d  (a = sc[6])
__global__ void k_line_step(double* __restrict__ x, const double* __restrict__ d,
                            double* __restrict__ r, const double* __restrict__ Sd,
                            const double* __restrict__ sc, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double a = sc[6];
  x[t] = x[t] + a * d[t];
  r[t] = r[t] - a * Sd[t];
}

// ---------------------------------------------------------------------------
// History-projection start: block partials of several dot products against
// one vector, their fixed-order sums, linear combinations and the scaling of
// a Gram-Schmidt column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mdot(VecSet A, const double* __restrict__ q, size_t n,
                                                 double* __restrict__ part) {
  __shared__ double sh[kHist + 1][kBlock / 64];
  double s[kHist + 1];
#pragma unroll
  for (int j = 0; j <= kHist; j++) s[j] = 0.0;
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)kDotBlocks * kBlock) {
    const double qt = q[t];
#pragma unroll
    for (int j = 0; j <= kHist; j++)
      if (j < A.k) s[j] += A.v[j][t] * qt;
  }
#pragma unroll
  for (int j = 0; j <= kHist; j++) {
    if (j < A.k) {
      const double r = block_sum(s[j], sh[j]);
      if (threadIdx.x == 0) part[j * kDotBlocks + blockIdx.x] = r;
    }
  }
}

__global__ void k_mreduce(const double* __restrict__ part, int k, double* __restrict__ coef) {
  const int j = threadIdx.x;
  if (j < k) coef[j] = sum_part(part, j);
}

// qa = (set ? 0 : qa) + sgn * sum_j coef[j] A_j ; xb likewise with B_j
__global__ void k_combine(VecSet A, VecSet B, const double* __restrict__ coef, double sgn, int set,
                          double* __restrict__ qa, double* __restrict__ xb, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a = 0.0, b = 0.0;
  for (int j = 0; j < A.k; j++) {
    const double cj = sgn * coef[j];
    if (qa) a += cj * A.v[j][t];
    if (xb) b += cj * B.v[j][t];
  }
  if (qa) qa[t] = set ? a : qa[t] + a;
  if (xb) xb[t] = set ? b : xb[t] + b;
}

// q, x /= ||q|| when ||q||^2 > kHistDrop^2 * raw2 (raw2: the column's squared
// norm before orthogonalisation), else q = x = 0 (a dependent column).
__global__ void k_hscale(double* __restrict__ q, double* __restrict__ x,
                         const double* __restrict__ nrm2, const double* __restrict__ raw2, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double v = *nrm2;
  const bool keep = v > 0.0 && v > (kHistDrop * kHistDrop) * (*raw2);
  const double inv = keep ? 1.0 / sqrt(v) : 0.0;
  q[t] = q[t] * inv;
  x[t] = x[t] * inv;
}

// check_slip, src/monitor.f90:223-231
__global__ __launch_bounds__(kBlock) void k_max_slip(DevGrid g, Markers m, const double* __restrict__ U,
                                                     const double* __restrict__ V,
                                                     const double* __restrict__ W,
                                                     const double* __restrict__ E1nH,
                                                     const double* __restrict__ fb,
                                                     unsigned long long* red) {
  const int l = wave_marker(m.nb);
  const int lane = threadIdx.x & 63;
  if (l >= m.nb) return;
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;
  double r[3];
  gather3<false>(m, l, lane, [&](int c, const Pt& q) {
    const double f = c == 0 ? U[fidx(q.X, q.Y, q.Z, nx, nyg)]
                            : (c == 1 ? V[fidx(q.X, q.Y, q.Z, nxg, ny)] : W[fidx(q.X, q.Y, q.Z, nxg, nyg)]);
    return q.w * f;
  }, r);
  double mx = 0.0;
  for (int c = 0; c < 3; c++) {
    const int o = c * m.nb + l;
    mx = fmax(mx, fabs(r[c] - E1nH[o] * fb[o] - m.ub[o]));
  }
  if (lane == 0) atomicMax(red, (unsigned long long)__double_as_longlong(mx));
}

inline dim3 blocks_for(size_t n, int b = kBlock) { return dim3((unsigned)((n + b - 1) / b)); }
inline dim3 wave_blocks(int nb) { return dim3((unsigned)((nb + kBlock / 64 - 1) / (kBlock / 64))); }

}  // namespace

hipError_t launch_setup(const DevGrid& g, const SetupGrid& sg, const Markers& m, int* bad,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_setup, blocks_for((size_t)4 * m.nb), dim3(kBlock), 0, s, g, sg, m, bad);
  return hipGetLastError();
}

hipError_t launch_wave_geometry(const WaveBody& w, const WaveTime& wt, double* xb, double* yb,
                                double* zb, double* normals, double* sb, double* ub, int* yref,
                                int* bad, hipStream_t s) {
  const size_t nb = 2 * (size_t)w.nxb * w.nzb;
  hipLaunchKernelGGL(k_wave_geometry, blocks_for(nb), dim3(kBlock), 0, s, w, wt, xb, yb, zb,
                     normals, sb, ub, yref, bad);
  return hipGetLastError();
}

hipError_t launch_interp_uvw(const DevGrid& g, const Markers& m, const double* U, const double* V,
                             const double* W, bool with_dxn, double* out, hipStream_t s) {
  if (with_dxn)
    hipLaunchKernelGGL(k_interp_uvw<true>, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, U, V, W, out);
  else
    hipLaunchKernelGGL(k_interp_uvw<false>, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, U, V, W, out);
  return hipGetLastError();
}

hipError_t launch_interp_c(const DevGrid& g, const Markers& m, const double* P, double sign,
                           double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_interp_c, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, P, sign, out);
  return hipGetLastError();
}

hipError_t launch_schur_rows(const DevGrid& g, const Markers& m, const double* phi, double dt,
                             const double* Ureg, const double* Vreg, const double* Wreg,
                             const double* E1nH, const double* dudn, double* Sf, hipStream_t s) {
  hipLaunchKernelGGL(k_schur_rows, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, phi, dt, Ureg, Vreg,
                     Wreg, E1nH, dudn, Sf);
  return hipGetLastError();
}

hipError_t launch_rhs_u(const DevGrid& g, const Markers& m, const double* U, const double* V,
                        const double* W, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_rhs_u, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, U, V, W, out);
  return hipGetLastError();
}

hipError_t launch_aux(const Markers& m, const double* f, double dt, double nu, double* aux,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_aux, blocks_for((size_t)3 * m.nb), dim3(kBlock), 0, s, m, f, dt, nu, aux);
  return hipGetLastError();
}

hipError_t launch_spread3(const Markers& m, const Spread sp[3], const double* f, double* F0,
                          double* F1, double* F2, hipStream_t s) {
  int mx = 0;
  for (int c = 0; c < 3; c++) mx = sp[c].nnodes > mx ? sp[c].nnodes : mx;
  if (mx == 0) return hipSuccess;
  dim3 grid = blocks_for((size_t)mx);
  grid.y = 3;
  hipLaunchKernelGGL(k_spread3, grid, dim3(kBlock), 0, s, sp[0], sp[1], sp[2], f, m.nb, F0, F1, F2);
  return hipGetLastError();
}

hipError_t launch_spread_coeffs(const Markers& m, const Spread& sp, int comp, hipStream_t s) {
  if (sp.nentries == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spread_coeffs, blocks_for(sp.nentries), dim3(kBlock), 0, s, m.nb, m.A,
                     m.factor, (const double*)m.st[comp].phi, m.sb, sp);
  return hipGetLastError();
}

hipError_t launch_spread(const Markers& m, const Spread& sp, int comp, const double* f, double* F,
                         hipStream_t s) {
  (void)m;
  (void)comp;
  if (sp.nnodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spread, blocks_for(sp.nnodes), dim3(kBlock), 0, s, sp, f, F);
  return hipGetLastError();
}

hipError_t launch_remove_mean(const Markers& m, double* f, const double* coef, const double* pj,
                              double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_body_sums, dim3(kMeanBlocks), dim3(kBlock), 0, s, m, (const double*)f, part);
  hipLaunchKernelGGL(k_remove_mean, blocks_for(m.nb), dim3(kBlock), 0, s, m, f, (const double*)part,
                     coef, pj);
  return hipGetLastError();
}

hipError_t launch_ib_correct(const DevGrid& g, const double* Uint, const double* Vint,
                             const double* Wint, const double* Ureg, const double* Vreg,
                             const double* Wreg, const double* Pint, const double* phi, double dt,
                             double* U, double* V, double* W, hipStream_t s) {
  hipLaunchKernelGGL(k_ib_correct, dim3((g.nxg + 63) / 64, (g.nyg + 3) / 4, g.nzg), dim3(kBlock), 0,
                     s, g, Uint, Vint, Wint, Ureg, Vreg, Wreg, Pint, phi, dt, U, V, W);
  return hipGetLastError();
}

hipError_t launch_affine(const double* x, double a, double b, double* y, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_affine, blocks_for(n), dim3(kBlock), 0, s, x, a, b, y, n);
  return hipGetLastError();
}

hipError_t launch_h_interp(const DevGrid& g, const double* Hc, double* Hu, double* Hv, double* Hw,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_h_interp, dim3((g.nxg - 2 + 63) / 64, (g.nyg - 2 + 3) / 4, g.nzg - 1),
                     dim3(kBlock), 0, s, g, Hc, Hu, Hv, Hw);
  return hipGetLastError();
}

hipError_t launch_dot(const double* a, const double* b, const double* a2, const double* b2,
                      size_t n, double* part, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(kDotBlocks), dim3(kBlock), 0, s, a, b, a2, b2, n, part, slot);
  return hipGetLastError();
}

hipError_t launch_norm_scales(const double* part, int sr, size_t n_r, int sf, size_t n_f,
                              double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_scales, dim3(1), dim3(64), 0, s, part, sr, (double)n_r, sf, (double)n_f,
                     out);
  return hipGetLastError();
}

hipError_t launch_bicg_scalar(int step, double* sc, const double* part, double* host_err,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_scalar, dim3(1), dim3(64), 0, s, step, sc, part, host_err);
  return hipGetLastError();
}

hipError_t launch_bicg_p(const double* r, const double* v, double* sc, const double* part, double* p,
                         size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_p, blocks_for(n), dim3(kBlock), 0, s, r, v, sc, part, p, n);
  return hipGetLastError();
}

hipError_t launch_bicg_hs(const double* x, const double* p, const double* r, const double* v,
                          double* sc, const double* part, double* h, double* sv, size_t n,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_hs, blocks_for(n), dim3(kBlock), 0, s, x, p, r, v, sc, part, h, sv, n);
  return hipGetLastError();
}

hipError_t launch_bicg_xr(const double* h, const double* sv, const double* t, const double* rhat,
                          double* sc, double* part, double* x, double* r, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_xr, dim3(kDotBlocks), dim3(kBlock), 0, s, h, sv, t, rhat, sc, part, x, r,
                     n);
  return hipGetLastError();
}

hipError_t launch_bicg_init(const double* b, double* r, double* rhat, double* p, double* v,
                            size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_init, blocks_for(n), dim3(kBlock), 0, s, b, r, rhat, p, v, n);
  return hipGetLastError();
}

hipError_t launch_line_step(double* x, const double* d, double* r, const double* Sd,
                            const double* sc, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_line_step, blocks_for(n), dim3(kBlock), 0, s, x, d, r, Sd, sc, n);
  return hipGetLastError();
}

hipError_t launch_mdot(const VecSet& A, const double* q, size_t n, double* part, hipStream_t s) {
  if (A.k < 1 || A.k > kHist + 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mdot, dim3(kDotBlocks), dim3(kBlock), 0, s, A, q, n, part);
  return hipGetLastError();
}

hipError_t launch_mreduce(const double* part, int k, double* coef, hipStream_t s) {
  if (k < 1 || k > kHist + 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mreduce, dim3(1), dim3(64), 0, s, part, k, coef);
  return hipGetLastError();
}

hipError_t launch_combine(const VecSet& A, const VecSet& B, const double* coef, double sgn, bool set,
                          double* qa, double* xb, size_t n, hipStream_t s) {
  if (A.k < 0 || A.k > kHist + 1 || (xb && B.k != A.k)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_combine, blocks_for(n), dim3(kBlock), 0, s, A, B, coef, sgn, set ? 1 : 0, qa,
                     xb, n);
  return hipGetLastError();
}

hipError_t launch_hscale(double* q, double* x, const double* nrm2, const double* raw2, size_t n,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_hscale, blocks_for(n), dim3(kBlock), 0, s, q, x, nrm2, raw2, n);
  return hipGetLastError();
}

hipError_t launch_max_slip(const DevGrid& g, const Markers& m, const double* U, const double* V,
                           const double* W, const double* E1nH, const double* fb,
                           unsigned long long* red, hipStream_t s) {
  hipLaunchKernelGGL(k_max_slip, wave_blocks(m.nb), dim3(kBlock), 0, s, g, m, U, V, W, E1nH, fb, red);
  return hipGetLastError();
}

}  // namespace ib
}  // namespace c3d
