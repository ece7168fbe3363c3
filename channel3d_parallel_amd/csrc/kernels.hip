// Projection-path kernels for gfx950 (FP64 plain path).
//
// Spectral layout: the half spectrum S is stored [j][kx][kz] with kz
// fastest (nym x nxh x nzp complex), so that
//   * the x transform writes 16-wide kz runs (LDS transpose),
//   * the z transform reads one contiguous column per (j, kx),
//   * the y tridiagonal runs one lane per (kx, kz) with lanes along kz,
//     coalesced at every j.
// Physical fields keep the reference's Fortran layout (x fastest,
// ghost-inclusive), so host copies need no reshaping.
#include <cstdlib>

#include "fft_lds.hpp"
#include "kernels.hpp"

namespace c3d {

namespace {

constexpr int kBlock = 256;
// z planes per lane in the gradient correction: the single potential, and
// the two-part potential of the refinement (which loads twice the potential
// words per face, so reusing p(k+1) over more planes pays)
constexpr int kGradKCDefault = 2;
constexpr int kGradKCSplitDefault = 4;

// Rows of the z tile handled by one x-transform workgroup (row pairs are
// packed into one complex transform).  Default: 16 rows (256-byte kz runs in
// the spectrum) up to x length 192, 8 rows above (LDS per workgroup stays
// near 28 KB so that several workgroups share a CU).  C3D_XTILE overrides it
// (tuning knob; must divide nzp).
int tile_rows(int nxp, int nzp) {
  static const int forced = [] {
    const char* e = getenv("C3D_XTILE");
    return e ? atoi(e) : 0;
  }();
  if ((forced == 2 || forced == 4 || forced == 8 || forced == 16) && nzp % forced == 0) return forced;
  const int want = nxp > 192 ? 8 : 16;
  for (int t = want; t >= 2; t /= 2)
    if (nzp % t == 0) return t;
  return 2;
}

// ---------------------------------------------------------------------------
// Forward x transform: div (or rhs) rows -> R2C half spectrum.
// Block (kBlock) handles y-plane j = blockIdx.y + 2 and z rows
// k = blockIdx.x*TK + 2 .. +TK-1.  Rows 2p and 2p+1 are transformed together
// as the real and imaginary parts of one complex row and separated after.
// ---------------------------------------------------------------------------
// Input sources of the forward x transform.
constexpr int kSrcRhs = 0;    // rhs_p-shaped FP64 array
constexpr int kSrcUVW = 1;    // divergence of U, V, W computed on the fly
constexpr int kSrcUVWB = 2;   // same, and the FP64 divergence also stored to bout
constexpr int kSrcUVWc = 4;   // as kSrcUVW with cached U and W loads (fields that fit in cache)
constexpr int kSrcResid = 5;  // FP64 residual b - M phi of the refinement, on the fly

template <typename real, int N, int TK, int SRC>
__global__ __launch_bounds__(kBlock) void k_fwd_x(DevGrid g, const double* __restrict__ U,
                                                  const double* __restrict__ V,
                                                  const double* __restrict__ W,
                                                  const void* __restrict__ rhs_in,
                                                  const cpx<real>* __restrict__ tw,
                                                  cpx<real>* __restrict__ S,
                                                  double* __restrict__ bout,
                                                  const double* __restrict__ gauge,
                                                  unsigned long long* __restrict__ rmax) {
  constexpr bool FROM_UVW = (SRC == kSrcUVW || SRC == kSrcUVWB || SRC == kSrcUVWc);
  constexpr bool RESID = (SRC == kSrcResid);
  constexpr bool NTL = (SRC != kSrcUVWc);  // non-temporal U and W loads
  constexpr int NV = FROM_UVW ? 5 : (RESID ? 6 : 1);  // values loaded per point
  constexpr int NP = TK / 2;
  constexpr int NT = kBlock / NP;
  constexpr int NH = N / 2 + 1;
  constexpr int LP = lds_row_len(N);  // padded FFT row
  constexpr int TKP = TK + 1;         // padded pitch of the [kx][row] layout
  constexpr int BUF = (NP * LP > NH * TKP) ? NP * LP : NH * TKP;
  __shared__ cpx<real> buf[BUF];

  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TK;
  const int jj = blockIdx.y;
  const int j = jj + 2;
  const int nx = g.nx, ny = g.ny, nxg = g.nxg, nyg = g.nyg;

  // 1) rows -> LDS.  Rows are walked with uniform (scalar) row bases; each
  //    lane owns the x positions n = tid, tid + kBlock, ...  Loads of a
  //    batch of RB rows are all issued before any is used.
  constexpr int NPL = (N + kBlock - 1) / kBlock;  // x positions per lane
  // rows per load batch (the residual loads 6 values per point: 2 rows)
  constexpr int RB = TK < 4 ? TK : (RESID ? 2 : 4);
  const double ry = g.rdyf[j];
  double wprev[NPL];  // W(k-1) of the next row: the W row of the row before
  if constexpr (FROM_UVW) {
    const double* W0 = W + fidx(1, j, k0 + 1, nxg, nyg);
#pragma unroll
    for (int p = 0; p < NPL; p++) {
      const int n = tid + p * kBlock;
      wprev[p] = (N % kBlock == 0 || n < N) ? W0[n + 1] : 0.0;
    }
  }
  // Residual r = b - M phi (D G on the unique cells, periodic in x and z, no
  // gradient on the wall faces, plus the (0,0)-mode gauge row at j = 2), with
  // U = phi and rhs_in = b; each lane walks its x positions down the tile's
  // rows carrying phi(k-1), phi(k).  bout (may be null) receives the FP64
  // residual.
  const double* phi = U;
  const size_t pk = (size_t)(nxg - 2) * (nyg - 2);  // rhs_p plane stride
  const int rowj = nxg - 2;                          // rhs_p row stride
  double pm[NPL], p0[NPL];
  // per-lane spacing reciprocals, loaded once for the whole tile (lanes past
  // the row end read row position N-1's and do not use them)
  double rxa[NPL], cxp[NPL], cxm[NPL];
#pragma unroll
  for (int p = 0; p < NPL; p++) {
    const int i = min(tid + p * kBlock, N - 1) + 2;
    rxa[p] = g.rdxf[i];
    cxp[p] = cxm[p] = 0.0;
    if constexpr (RESID) {
      cxp[p] = g.rdxg[i];
      cxm[p] = g.rdxg[(i == 2) ? nx - 1 : i - 1];
    }
  }
  double amax = 0.0;
  double gj = 0.0, rhy = 0.0, rhym = 0.0;
  // the y neighbours' row offsets: on the first and last rows (no gradient
  // on the wall face) the row itself, read with a zero coefficient, so the
  // loads need no branch
  int jpo = 0, jmo = 0;
  if constexpr (RESID) {
    gj = (j == 2) ? *gauge : 0.0;
    rhy = (j <= ny - 1) ? g.rdyg[j] : 0.0;
    rhym = (j >= 3) ? g.rdyg[j - 1] : 0.0;
    jpo = (j <= ny - 1) ? rowj : 0;
    jmo = (j >= 3) ? rowj : 0;
    const int kf = k0 + 2, km0 = (kf == 2) ? g.nz - 1 : kf - 1;
#pragma unroll
    for (int p = 0; p < NPL; p++) {
      const int n = tid + p * kBlock;
      pm[p] = p0[p] = 0.0;
      if (N % kBlock == 0 || n < N) {
        pm[p] = phi[ridx(2, j, km0, nxg, nyg) + n];
        p0[p] = phi[ridx(2, j, kf, nxg, nyg) + n];
      }
    }
  }
#pragma unroll
  for (int t0 = 0; t0 < TK; t0 += RB) {
    double v[RB][NPL][NV];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int k = k0 + t0 + r + 2;
      if constexpr (FROM_UVW) {
        const double* Ur = U + fidx(1, j, k, nx, nyg);
        const double* Vr = V + fidx(1, j, k, nxg, ny);
        const double* Wr = W + fidx(1, j, k, nxg, nyg);
        const double* Vm = Vr - nxg;
#pragma unroll
        for (int p = 0; p < NPL; p++) {
          const int n = tid + p * kBlock;
          if (N % kBlock == 0 || n < N) {
            // x index i = n + 2 -> offset n + 1 in the row.  U and W rows
            // are read by this workgroup only: non-temporal when the fields
            // are too large to stay in cache between kernels.  V rows are
            // read twice (as V(j) here and V(j-1) by the next y row's
            // workgroup): cached.
            if constexpr (NTL) {
              v[r][p][0] = __builtin_nontemporal_load(&Ur[n + 1]);
              v[r][p][1] = __builtin_nontemporal_load(&Ur[n]);
              v[r][p][4] = __builtin_nontemporal_load(&Wr[n + 1]);
            } else {
              v[r][p][0] = Ur[n + 1];
              v[r][p][1] = Ur[n];
              v[r][p][4] = Wr[n + 1];
            }
            v[r][p][2] = Vr[n + 1];
            v[r][p][3] = Vm[n + 1];
          }
        }
      } else if constexpr (RESID) {
        const double* Br = static_cast<const double*>(rhs_in) + ridx(2, j, k, nxg, nyg);
        const double* Pr = phi + ridx(2, j, k, nxg, nyg);
#pragma unroll
        for (int p = 0; p < NPL; p++) {
          const int n = tid + p * kBlock;
          if (N % kBlock == 0 || n < N) {
            v[r][p][0] = Br[n];
            v[r][p][1] = Pr[n + pk];                          // phi(i, j, k+1)
            v[r][p][2] = Pr[n + 1];                           // phi(i+1, j, k)
            v[r][p][3] = Pr[(n == 0) ? N - 1 : n - 1];  // phi(i-1), periodic at i = 2
            v[r][p][4] = Pr[n + jpo];                   // phi(i, j+1, k)
            v[r][p][5] = Pr[n - jmo];                   // phi(i, j-1, k)
          }
        }
      } else {
        const double* Rr = static_cast<const double*>(rhs_in) + ridx(2, j, k, nxg, nyg);
#pragma unroll
        for (int p = 0; p < NPL; p++) {
          const int n = tid + p * kBlock;
          if (N % kBlock == 0 || n < N) v[r][p][0] = Rr[n];
        }
      }
    }
#pragma unroll
    for (int p = 0; p < NPL; p++) {
      const int n = tid + p * kBlock;
      if (N % kBlock == 0 || n < N) {
        const double rx = rxa[p];
#pragma unroll
        for (int r = 0; r < RB; r += 2) {  // rows 2q, 2q+1 -> one complex
          double d[2];
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if constexpr (FROM_UVW) {
              // divergence, src/differential_operators.f90:64-72, with the
              // face-spacing reciprocals tabulated
              const double rz = g.rdzf[k0 + t0 + r + h + 2];
              const double wm = (r + h == 0) ? wprev[p] : v[r + h - 1][p][4];
              d[h] = (v[r + h][p][0] - v[r + h][p][1]) * rx +
                     (v[r + h][p][2] - v[r + h][p][3]) * ry + (v[r + h][p][4] - wm) * rz;
            } else if constexpr (RESID) {
              // r = b - D G phi on the unique cells (i = 2..nxp+1,
              // j = 2..nyg-1, k = 2..nzp+1): periodic in x and z (the face
              // left of cell 2 is the image of face nx-1, resp. nz-1), no
              // gradient on the wall faces (j = 1, ny); phi carries its
              // periodic duplicates (i = nxg-1, k = nzp+2)
              const int i = n + 2, k = k0 + t0 + r + h + 2;
              const int km = (k == 2) ? g.nz - 1 : k - 1;
              const double* q = v[r + h][p];
              const double c0 = p0[p];
              const double gxp = (q[2] - c0) * cxp[p];
              const double gxm = (c0 - q[3]) * cxm[p];
              const double gyp = (q[4] - c0) * rhy;
              const double gym = (c0 - q[5]) * rhym;
              const double gzp = (q[1] - c0) * g.rdzg[k];
              const double gzm = (c0 - pm[p]) * g.rdzg[km];
              const double lap = (gxp - gxm) * rx + (gyp - gym) * ry + (gzp - gzm) * g.rdzf[k];
              d[h] = (q[0] - lap) - gj;
              amax = fmax(amax, fabs(d[h]));
              pm[p] = c0;
              p0[p] = q[1];
              // the FP64 residual itself, when later sweeps start from it
              if (bout) bout[ridx(i, j, k, nxg, nyg)] = d[h];
              // an FP32 transform takes the residual rounded to FP32
              if constexpr (sizeof(real) == 4) d[h] = (double)(float)d[h];
            } else {
              d[h] = v[r + h][p][0];
            }
            if constexpr (SRC == kSrcUVWB) bout[ridx(n + 2, j, k0 + t0 + r + h + 2, nxg, nyg)] = d[h];
          }
          buf[((t0 + r) >> 1) * LP + lp(n)] = cmk((real)d[0], (real)d[1]);
        }
        if constexpr (FROM_UVW) wprev[p] = v[RB - 1][p][4];
      }
    }
  }
  __syncthreads();

  // 2) complex FFT of each packed pair
  fft_lds<real, N, NT, -1>(buf + (tid / NT) * LP, tw, tid % NT);

  // 3) separate the two real transforms: A_k = (Z_k + conj Z_{N-k})/2,
  //    B_k = (Z_k - conj Z_{N-k})/(2i); store transposed [kx][row].
  constexpr int PER = (NP * NH + kBlock - 1) / kBlock;
  cpx<real> A[PER], B[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * kBlock;
    if (e < NP * NH) {
      const int p = e / NH, kx = e % NH;
      const cpx<real> z = buf[p * LP + lp(kx)];
      const cpx<real> zc = buf[p * LP + lp(kx == 0 ? 0 : N - kx)];
      A[q] = cmk((real)0.5 * (z.x + zc.x), (real)0.5 * (z.y - zc.y));
      B[q] = cmk((real)0.5 * (z.y + zc.y), (real)0.5 * (zc.x - z.x));
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * kBlock;
    if (e < NP * NH) {
      const int p = e / NH, kx = e % NH;
      buf[kx * TKP + 2 * p] = A[q];
      buf[kx * TKP + 2 * p + 1] = B[q];
    }
  }
  __syncthreads();

  // 4) LDS -> S[j][kx][k0 .. k0+TK)
  const int nzp = g.nzp;
  for (int e = tid; e < NH * TK; e += kBlock) {
    const int kx = e / TK, t = e % TK;
    // streaming store: the spectrum is next read by another kernel
    const cpx<real> v = buf[kx * TKP + t];
    cpx<real>* d = &S[((size_t)jj * NH + kx) * nzp + k0 + t];
    __builtin_nontemporal_store(v.x, &d->x);
    __builtin_nontemporal_store(v.y, &d->y);
  }
  if constexpr (RESID) {
    // max |r| of the workgroup, one atomic per workgroup into one of 64
    // slots (linear workgroup index % 64) so that the workgroups do not
    // serialise on one address; the host takes the max of the slots
    __shared__ double sm[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, 64));
    if ((tid & 63) == 0) sm[tid >> 6] = amax;
    __syncthreads();
    if (tid == 0 && rmax) {
      double mx = sm[0];
      for (int w = 1; w < kBlock / 64; w++) mx = fmax(mx, sm[w]);
      const unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
      atomicMax(rmax + (lin & 63), (unsigned long long)__double_as_longlong(mx));
    }
  }
}

// ---------------------------------------------------------------------------
// z transform (C2C) of every column S[j][kx][:], in place.
// ---------------------------------------------------------------------------
template <typename real, int N, int SIGN, int NT>
__global__ __launch_bounds__(kBlock) void k_fft_z(cpx<real>* __restrict__ S,
                                                  const cpx<real>* __restrict__ tw, int ncol) {
  // NT threads per column, NC columns per workgroup
  constexpr int NC = kBlock / NT;
  constexpr int LP = lds_row_len(N);
  __shared__ cpx<real> buf[NC * LP];
  const int tid = threadIdx.x;
  const size_t col0 = (size_t)blockIdx.x * NC;
  {
    // every lane loads (element 0 in place of one past the end) and keeps
    // zero where it has no element: no branch per load, so that all ITEMS
    // loads are in flight together (behind branches, the FP32 ones each
    // waited for their data)
    constexpr int ITEMS = (NC * N + kBlock - 1) / kBlock;
    real vx[ITEMS], vy[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; q++) {
      const int e = tid + q * kBlock;
      const bool ok = e < NC * N && col0 + e / N < (size_t)ncol;
      const cpx<real> t = S[ok ? col0 * N + e : 0];
      vx[q] = ok ? t.x : (real)0;
      vy[q] = ok ? t.y : (real)0;
    }
#pragma unroll
    for (int q = 0; q < ITEMS; q++) {
      const int e = tid + q * kBlock;
      if (e < NC * N) buf[(e / N) * LP + lp(e % N)] = cmk(vx[q], vy[q]);
    }
  }
  __syncthreads();
  fft_lds<real, N, NT, SIGN>(buf + (tid / NT) * LP, tw, tid % NT);
  for (int e = tid; e < NC * N; e += kBlock) {
    const int c = e / N;
    if (col0 + c < (size_t)ncol) S[col0 * N + e] = buf[c * LP + lp(e % N)];
  }
}

// ---------------------------------------------------------------------------
// Tridiagonal solve along y for every mode (kx, kz): src/poisson.f90:50-75.
// Matrix: D(j) = (Dyy(j,j) + kxx) + kzz, DL(j) = Dyy(j+1,j), DU(j) = Dyy(j,j+1),
// and D(2) *= 3/2 for the (0,0) mode.  Thomas elimination without
// pivoting (rows are diagonally dominant; agrees with partial pivoting to
// round-off).  One lane per mode, lanes along kz; CP holds the forward
// multipliers for the back substitution.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_tri64(DevGrid g, cpx<double>* __restrict__ S,
                                                  double* __restrict__ CP) {
  const int nzp = g.nzp, nym = g.nym;
  const size_t stride = (size_t)g.nxh * nzp;
  const size_t m = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= stride) return;
  const int kx = (int)(m / nzp), kz = (int)(m % nzp);
  const double lamx = g.kxx[kx], lamz = g.kzz[kz];
  cpx<double>* s = S + m;
  double* cp = CP + m;

  double d0 = (g.dyy_d[0] + lamx) + lamz;
  if (kx == 0 && kz == 0) d0 = 1.5 * d0;
  cpx<double> bprev = s[0];
  double den = d0;
  double cprev = (nym > 1) ? g.dyy_u[0] / den : 0.0;
  bprev = cmk(bprev.x / den, bprev.y / den);
  s[0] = bprev;
  cp[0] = cprev;
  for (int jj = 1; jj < nym; jj++) {
    const double l = g.dyy_l[jj - 1];
    den = ((g.dyy_d[jj] + lamx) + lamz) - l * cprev;
    cpx<double> b = s[(size_t)jj * stride];
    b = cmk((b.x - l * bprev.x) / den, (b.y - l * bprev.y) / den);
    s[(size_t)jj * stride] = b;
    if (jj < nym - 1) {
      cprev = g.dyy_u[jj] / den;
      cp[(size_t)jj * stride] = cprev;
    }
    bprev = b;
  }
  cpx<double> xnext = bprev;
  for (int jj = nym - 2; jj >= 0; jj--) {
    const double c = cp[(size_t)jj * stride];
    cpx<double> b = s[(size_t)jj * stride];
    b = cmk(b.x - c * xnext.x, b.y - c * xnext.y);
    s[(size_t)jj * stride] = b;
    xnext = b;
  }
}

// ---------------------------------------------------------------------------
// Same system for NYM = 8*M rows (64, 128, 192), partitioned: each mode's rows are split into 8
// parts of M rows, one lane per part, so every row is read once and written
// once and the whole part stays in registers.  A workgroup holds 32 modes;
// each wave holds 16 of them for 4 of the parts (lane = part*16 + mode), so
// that each row access is a run of 16 consecutive modes (256 bytes in FP64,
// 128 in FP32), and the two waves of a mode group exchange interface rows
// through LDS.
//   pass 1  forward elimination inside the part with the unknown x_{a-1}
//           above it carried as a spike column sp:
//             x_i = y_i - sp_i x_{a-1} - cp_i x_{i+1}
//   pass 2  back-substituted forms of the part's first and last rows,
//             x_i = G_i - V_i x_{a-1} - W_i x_b   (i = a, b-1)
//   reduced the 2x8 interface equations of a mode, solved by its part-0
//           lane through LDS: L_q = x(first row of q), R_q = x(last row);
//           forward L_q = lam_q + mu_q L_{q+1}, R_q = rho_q + sig_q L_{q+1},
//           back L_q, then R_{q-1} = rho_{q-1} + sig_{q-1} L_q
//   pass 3  back substitution with the boundary values x_{a-1} = R_{q-1},
//           x_b = L_{q+1}.
// Part 0 (l = 0 in its first row) and the last part (u = 0 in its last row)
// have no outside coupling, so the reduced system has no boundary terms.
// ---------------------------------------------------------------------------
// Mode selection of the partitioned solve (mixed precision runs the solve
// in two launches: FP32 arithmetic for the well-conditioned modes, FP64 for
// the (0,0) mode and the lowest |k|, whose systems are nearly singular).
constexpr int kTriAll = 0;   // every mode
constexpr int kTriHigh = 1;  // modes with |kxx + kzz| >= lam_thr
constexpr int kTriList = 2;  // the modes list[0 .. nlist)

// real: storage of the right-hand side read from S; rout: storage of the
// solution written to Sout (in place when both are the same array).
// P: parts per mode; MWX (when set) overrides the modes per wave.
template <typename real, typename acc, int M, int SEL, typename rout = real, int P = 8, int MWX = 0>
__global__ __launch_bounds__(kBlock) void k_tri_part(DevGrid g, cpx<real>* S,
                                                     const int* __restrict__ list, int nlist,
                                                     double lam_thr, cpx<rout>* Sout) {
  constexpr int N = P * M;        // rows
  // modes per wave: a row access of a wave is MW consecutive modes, 256
  // contiguous bytes of the input spectrum (16 FP64 or 32 FP32 complex)
  constexpr int MW = MWX ? MWX : 256 / (int)sizeof(cpx<real>);
  constexpr int PW = 64 / MW;     // parts per wave
  constexpr int WG = P / PW;      // waves that cover the parts of a mode group
  static_assert(WG * PW == P && kBlock / 64 % WG == 0, "waves must tile the parts");
  constexpr int MB = kBlock / P;  // modes per workgroup
  constexpr int CPT = M + 1;      // padded pitch of the row tables
  constexpr int RP = 8 * P + 1;   // padded pitch of a mode's interface rows
  __shared__ acc td[P * CPT], tl[P * CPT], tu[P * CPT];
  __shared__ acc red[MB * RP];
  const int tid = threadIdx.x;
  for (int e = tid; e < N; e += kBlock) {
    // three loads in flight at once: clamped in-bounds indices (dyy_l and
    // dyy_u hold N-1 entries), the end rows' zero couplings selected after
    const int q = e / M, r = e % M;
    const double d = g.dyy_d[e];
    const double l = g.dyy_l[e > 0 ? e - 1 : 0];
    const double u = g.dyy_u[e < N - 1 ? e : 0];
    td[q * CPT + r] = (acc)d;
    tl[q * CPT + r] = (e == 0) ? (acc)0 : (acc)l;
    tu[q * CPT + r] = (e == N - 1) ? (acc)0 : (acc)u;
  }
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int p = (wv % WG) * PW + lane / MW;
  const int mb = (wv / WG) * MW + lane % MW;
  const int nzp = g.nzp;
  const size_t stride = (size_t)g.nxh * nzp;
  bool valid;
  size_t mv;
  if constexpr (SEL == kTriList) {
    const int e = blockIdx.x * MB + mb;
    valid = e < nlist;
    mv = valid ? (size_t)list[e] : 0;
  } else {
    const size_t m = (size_t)blockIdx.x * MB + mb;
    valid = m < stride;
    mv = valid ? m : 0;
  }
  const int kx = (int)(mv / nzp), kz = (int)(mv % nzp);
  if constexpr (SEL == kTriHigh) valid = valid && !(fabs(g.kxx[kx] + g.kzz[kz]) < lam_thr);
  const acc lamx = (acc)g.kxx[kx];
  const acc lamz = (acc)g.kzz[kz];
  const bool zero_first = (kx == 0 && kz == 0 && p == 0);
  cpx<real>* s = S + mv + (size_t)(p * M) * stride;
  cpx<rout>* so = Sout + mv + (size_t)(p * M) * stride;

  // All M rows are loaded unconditionally and converted afterwards: mv is a
  // valid mode for every lane, so the addresses stay in bounds, and without
  // a branch per row (whose FP32 -> FP64 conversion waited for its load) the
  // M loads of a lane are in flight together.
  acc yr[M], yi[M], cp[M], sp[M];
  {
    cpx<real> t[M];
#pragma unroll
    for (int r = 0; r < M; r++) t[r] = s[(size_t)r * stride];
#pragma unroll
    for (int r = 0; r < M; r++) {
      yr[r] = valid ? (acc)t[r].x : (acc)0;
      yi[r] = valid ? (acc)t[r].y : (acc)0;
    }
  }
  __syncthreads();  // row tables

  // pass 1
  const acc* rd = td + p * CPT;
  const acc* rl = tl + p * CPT;
  const acc* ru = tu + p * CPT;
#pragma unroll
  for (int r = 0; r < M; r++) {
    acc d = (rd[r] + lamx) + lamz;
    const acc l = rl[r];
    if (r == 0) {
      if (zero_first) d = (acc)1.5 * d;
      const acc inv = (acc)1 / d;
      yr[0] = yr[0] * inv;
      yi[0] = yi[0] * inv;
      sp[0] = l * inv;
      cp[0] = ru[0] * inv;
    } else {
      const acc inv = (acc)1 / (d - l * cp[r - 1]);
      yr[r] = (yr[r] - l * yr[r - 1]) * inv;
      yi[r] = (yi[r] - l * yi[r - 1]) * inv;
      sp[r] = -(l * sp[r - 1]) * inv;
      cp[r] = ru[r] * inv;
    }
  }
  // pass 2
  acc* my = red + mb * RP + p * 8;
  {
    acc gr = yr[M - 1], gi = yi[M - 1], v = sp[M - 1], w = cp[M - 1];
    my[4] = gr;
    my[5] = gi;
    my[6] = v;
    my[7] = w;
#pragma unroll
    for (int r = M - 2; r >= 0; r--) {
      gr = yr[r] - cp[r] * gr;
      gi = yi[r] - cp[r] * gi;
      v = sp[r] - cp[r] * v;
      w = -(cp[r] * w);
    }
    my[0] = gr;
    my[1] = gi;
    my[2] = v;
    my[3] = w;
  }
  __syncthreads();
  // reduced system: slot q holds GL(2) VL WL GR(2) VR WR on entry; the
  // forward sweep leaves lam(2) mu sig_{q-1} rho_{q-1}(2) in it and the
  // back sweep x_{a-1}(2) x_b(2) in positions 4..7
  if (p == 0) {
    acc* rs = red + mb * RP;
    acc rhr = 0, rhi = 0, sig = 0;
    for (int q = 0; q < P; q++) {
      acc* e = rs + q * 8;
      const acc glr = e[0], gli = e[1], vl = e[2], wl = e[3];
      const acc grr = e[4], gri = e[5], vr = e[6], wr = e[7];
      const acc iden = (acc)1 / ((acc)1 + vl * sig);
      const acc lr = (glr - vl * rhr) * iden, li = (gli - vl * rhi) * iden;
      const acc mu = -wl * iden;
      e[0] = lr;
      e[1] = li;
      e[2] = mu;
      e[3] = sig;
      e[4] = rhr;
      e[5] = rhi;
      const acc nr = grr - vr * (rhr + sig * lr), ni = gri - vr * (rhi + sig * li);
      sig = -(vr * sig * mu + wr);
      rhr = nr;
      rhi = ni;
    }
    acc lnr = 0, lni = 0;  // L_{q+1}
    for (int q = P - 1; q >= 0; q--) {
      acc* e = rs + q * 8;
      const acc lr = e[0] + e[2] * lnr, li = e[1] + e[2] * lni;  // L_q
      const acc ar = e[4] + e[3] * lr, ai = e[5] + e[3] * li;     // R_{q-1}
      e[6] = lnr;
      e[7] = lni;
      e[4] = ar;
      e[5] = ai;
      lnr = lr;
      lni = li;
    }
  }
  __syncthreads();
  // pass 3
  const acc xar = my[4], xai = my[5];
  acc xnr = my[6], xni = my[7];
#pragma unroll
  for (int r = M - 1; r >= 0; r--) {
    const acc xr = (yr[r] - sp[r] * xar) - cp[r] * xnr;
    const acc xi = (yi[r] - sp[r] * xai) - cp[r] * xni;
    if (valid) so[(size_t)r * stride] = cmk((rout)xr, (rout)xi);
    xnr = xr;
    xni = xi;
  }
}

// ---------------------------------------------------------------------------
// Inverse x transform (C2R) of S[j][kx][k0..k0+TK) and the output stage of
// solve_poisson_equation (src/poisson.f90:78-99):
//   phi = Re(IDFT)/(nxp*nzp); periodic copies i = nxg-1 <- 2, k = nzp+2 <- 2.
// MODE 0: out = phi (rhs_p), MODE 1: out = phi/dt (P_interim, with the zero
// plane nzg that the divergence leaves), rk1: P = phi/(dt*8/15).
// ---------------------------------------------------------------------------
// Refinement modes, the potential kept as the first solve's phi1 plus a
// correction phic (their FP64 sum would round phi1's low bits away, and the
// rounding of a large smooth potential, amplified by D G ~ 1/dy^2, bounds
// max|div| far above the velocity's own rounding):
//   MODE 3: out = phi_in + phi (correction accumulated),
//   MODE 5: out2 = phic' = phi_in2 + phi (phi_in2 null: phi); rk1: P =
//           (phi_in + phic')/(dt*8/15).  P_interim = (phi1 + phic')/dt is
//           written by the split gradient correction, which reads both.
template <typename real, int N, int TK, int MODE>
__global__ __launch_bounds__(kBlock) void k_inv_x(DevGrid g, const cpx<real>* __restrict__ S,
                                                  const cpx<real>* __restrict__ tw,
                                                  double* out, double* __restrict__ P,
                                                  double dt, int rk1, const double* phi_in,
                                                  const double* phi_in2, double* out2) {
  constexpr int NP = TK / 2;
  constexpr int NT = kBlock / NP;
  constexpr int NH = N / 2 + 1;
  constexpr int LP = lds_row_len(N);
  constexpr int TKP = TK + 1;
  constexpr int BUF = (NP * LP > NH * TKP) ? NP * LP : NH * TKP;
  __shared__ cpx<real> buf[BUF];

  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * TK;
  const int jj = blockIdx.y;
  const int j = jj + 2;
  const int nzp = g.nzp, nxg = g.nxg, nyg = g.nyg;

  {
    constexpr int ITEMS = (NH * TK + kBlock - 1) / kBlock;
    real vx[ITEMS], vy[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; q++) {
      const int e = tid + q * kBlock;
      vx[q] = (real)0;
      vy[q] = (real)0;
      if (e < NH * TK) {
        const int kx = e / TK, t = e % TK;
        const cpx<real>* sp = &S[((size_t)jj * NH + kx) * nzp + k0 + t];
        cpx<real> v;
        v.x = __builtin_nontemporal_load(&sp->x);  // read once
        v.y = __builtin_nontemporal_load(&sp->y);
        vx[q] = v.x;
        vy[q] = v.y;
      }
    }
#pragma unroll
    for (int q = 0; q < ITEMS; q++) {
      const int e = tid + q * kBlock;
      if (e < NH * TK) buf[(e / TK) * TKP + e % TK] = cmk(vx[q], vy[q]);
    }
  }
  __syncthreads();

  // Hermitian extension and packing Z = A + iB (imaginary parts of the
  // DC and Nyquist bins dropped: the real part of the full inverse).
  constexpr int PER = (NP * N + kBlock - 1) / kBlock;
  cpx<real> Z[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * kBlock;
    if (e < NP * N) {
      const int p = e / N, n = e % N;
      cpx<real> a, b;
      if (n <= N / 2) {
        a = buf[n * TKP + 2 * p];
        b = buf[n * TKP + 2 * p + 1];
        if (n == 0 || n == N / 2) {
          a.y = (real)0;
          b.y = (real)0;
        }
      } else {
        a = buf[(N - n) * TKP + 2 * p];
        b = buf[(N - n) * TKP + 2 * p + 1];
        a.y = -a.y;
        b.y = -b.y;
      }
      Z[q] = cmk(a.x - b.y, a.y + b.x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * kBlock;
    if (e < NP * N) buf[(e / N) * LP + lp(e % N)] = Z[q];
  }
  __syncthreads();

  fft_lds<real, N, NT, +1>(buf + (tid / NT) * LP, tw, tid % NT);

  const double norm = (double)((long long)g.nxp * (long long)nzp);
  const double cP = dt * (8.0 / 15.0);
  const int idup = nxg - 1;     // x duplicate of i = 2
  const int kdup = nzp + 2;     // z duplicate of k = 2
  // rows walked with uniform row bases; lane owns x positions tid + p*kBlock
  constexpr int NPL = (N + kBlock - 1) / kBlock;
  cpx<real> zpair[NPL];
#pragma unroll
  for (int t = 0; t < TK; t++) {
    const int k = k0 + t + 2;
    double* orow = out + ridx(2, j, k, nxg, nyg);      // x index 2 at offset 0
    double* prow = P + fidx(2, j, k, nxg, nyg);
    const bool dupz = (k0 == 0 && t == 0);             // row k = 2 also fills k = nzp+2
    double* orow2 = out + ridx(2, j, kdup, nxg, nyg);
    double* prow2 = P + fidx(2, j, kdup, nxg, nyg);
#pragma unroll
    for (int p = 0; p < NPL; p++) {
      const int n = tid + p * kBlock;
      if (N % kBlock == 0 || n < N) {
        // rows 2q and 2q+1 share one complex; the even row reads it
        cpx<real> z;
        if ((t & 1) == 0) zpair[p] = buf[(t >> 1) * LP + lp(n)];
        z = zpair[p];
        double phi = (double)((t & 1) ? z.y : z.x) / norm;
        const size_t at = ridx(2, j, k, nxg, nyg) + n;
        double corr = 0.0;  // MODE 5: the new correction phic'
        if constexpr (MODE == 3) phi = phi_in[at] + phi;
        if constexpr (MODE == 5) {
          corr = phi_in2 ? phi_in2[at] + phi : phi;
          if (rk1) phi = phi_in[at] + corr;  // the potential of P
        }
        constexpr bool OUT = MODE != 5;  // MODE 5 writes phic' (and P), not out
        const double o = (MODE == 1) ? phi / dt : phi;
        if constexpr (OUT) __builtin_nontemporal_store(o, &orow[n]);
        if constexpr (MODE == 5) out2[at] = corr;
        if (rk1) prow[n] = phi / cP;
        if (n == 0) {  // x duplicate i = nxg-1 (offset nxp)
          if constexpr (OUT) orow[N] = o;
          if constexpr (MODE == 5) out2[at + N] = corr;
          if (rk1) prow[N] = phi / cP;
        }
        if (dupz) {
          const size_t at2 = ridx(2, j, kdup, nxg, nyg) + n;
          if constexpr (OUT) orow2[n] = o;
          if constexpr (MODE == 5) out2[at2] = corr;
          if (rk1) prow2[n] = phi / cP;
          if (n == 0) {
            if constexpr (OUT) orow2[N] = o;
            if constexpr (MODE == 5) out2[at2 + N] = corr;
            if (rk1) prow2[N] = phi / cP;
          }
        }
      }
    }
  }
  (void)idup;
  if (MODE == 1 && k0 == 0) {
    // rhs_p plane nzg is zero after the divergence; P_interim = 0/dt there.
    for (int i = 2 + tid; i <= nxg - 1; i += kBlock) out[ridx(i, j, g.nzg, nxg, nyg)] = 0.0;
  }
  if (MODE == 0 && k0 == 0 && tid == 0) {
    // plane nzg keeps its input, but the x-periodic copy (src/poisson.f90:110)
    // also covers it: rhs(nxg-1, j, nzg) = rhs(2, j, nzg).
    out[ridx(idup, j, g.nzg, nxg, nyg)] = out[ridx(2, j, g.nzg, nxg, nyg)];
  }
}

// gauge = 0.5 * Dyy(2,2) * mean over the unique cells of phi(:, 2, :): the
// spectral solver's (0,0) mode uses D(2)*3/2 (src/poisson.f90:69), i.e. the
// operator L + 0.5*Dyy(2,2)*(plane mean at j = 2); a residual of that
// operator makes refinement converge to the FP64 solver's gauge.  Row sums
// per z (one block each), then one block sums the rows: fixed order.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(kBlock) void k_gauge_rows(DevGrid g, const double* __restrict__ phi,
                                                       double* __restrict__ rows) {
  __shared__ double sh[kBlock / 64];
  const int k = blockIdx.x + 2;
  double acc = 0.0;
  for (int i = 2 + threadIdx.x; i <= g.nxp + 1; i += kBlock) acc += phi[ridx(i, 2, k, g.nxg, g.nyg)];
  const double t = block_sum256(acc, sh);
  if (threadIdx.x == 0) rows[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_gauge_final(DevGrid g, const double* __restrict__ rows,
                                                        double* __restrict__ gauge) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int k = threadIdx.x; k < g.nzp; k += kBlock) acc += rows[k];
  const double t = block_sum256(acc, sh);
  if (threadIdx.x == 0) *gauge = 0.5 * g.dyy_d[0] * (t / ((double)g.nxp * (double)g.nzp));
}

// ---------------------------------------------------------------------------
// P ghost copies (src/poisson.f90:93-98): x, then y, then z.
// ---------------------------------------------------------------------------
__global__ void k_pghost_x(DevGrid g, double* __restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
  if (j > g.nyg) return;
  P[fidx(1, j, k, g.nxg, g.nyg)] = P[fidx(2, j, k, g.nxg, g.nyg)];
  P[fidx(g.nxg, j, k, g.nxg, g.nyg)] = P[fidx(g.nxg - 1, j, k, g.nxg, g.nyg)];
}
__global__ void k_pghost_y(DevGrid g, double* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
  if (i > g.nxg) return;
  P[fidx(i, 1, k, g.nxg, g.nyg)] = P[fidx(i, 2, k, g.nxg, g.nyg)];
  P[fidx(i, g.nyg, k, g.nxg, g.nyg)] = P[fidx(i, g.nyg - 1, k, g.nxg, g.nyg)];
}
__global__ void k_pghost_z(DevGrid g, double* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
  if (i > g.nxg) return;
  P[fidx(i, j, 1, g.nxg, g.nyg)] = P[fidx(i, j, 2, g.nxg, g.nyg)];
  P[fidx(i, j, g.nzg, g.nxg, g.nyg)] = P[fidx(i, j, g.nzg - 1, g.nxg, g.nyg)];
}

// ---------------------------------------------------------------------------
// Gradient correction: U = U - G(dt*Pint) on the index ranges of
// src/differential_operators.f90:28-50 (faces outside them keep U).
// SPLIT: U = U - (G(phi1) + G(phic)) for a potential kept as two parts (the
// refinement's phi1 + phic); Pint is phi1 then, and each part's differences
// are formed before they are added, so phic is not rounded into phi1.  The
// split kernel also writes P_interim = (phi1 + phic)/dt to pout on the whole
// rhs_p shape (the duplicates i = nxg-1 from i = 2, the zero plane nzg).
// ---------------------------------------------------------------------------
template <int KC, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_grad_correct(DevGrid g, double* __restrict__ U,
                                                         double* __restrict__ V,
                                                         double* __restrict__ W,
                                                         const double* __restrict__ Pint,
                                                         const double* __restrict__ phic,
                                                         double* __restrict__ pout, double dt) {
  // one lane per (i, j), walking KC consecutive z planes so that
  // p(k+1) of one plane is p(k) of the next; XCD-aware tile order so the
  // p(j+1) halo row is usually in the same L2
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);
  const int i = bx * 64 + (threadIdx.x & 63) + 2;
  // one row per wavefront: j is wave-uniform (scalar row arithmetic, 1/dy
  // a scalar load)
  const int j = __builtin_amdgcn_readfirstlane((int)(by * 4 + (threadIdx.x >> 6) + 2));
  const int kb = bz * KC + 2;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  if (i > nx - 1 || j > nyg - 1) return;
  // the potential at i+1 is the next lane's (a wavefront holds 64
  // consecutive i of one row): moved across lanes, not loaded again.  The
  // wave's last face ie (lane 63, or i = nx-1 in a row's last wave) takes
  // it from a load at one address for the whole wave (one cache line).
  const int ie = min((int)bx * 64 + 65, nx - 1);
  const double hx = g.dxg[i];
  const double hy = (j <= ny - 1) ? g.dyg[j] : 1.0;
  // potential at a cell: dt*Pint (the reference's operand), or the two parts
  auto pot = [&](size_t at, double& c) -> double {
    if constexpr (SPLIT) {
      c = phic[at];
      return Pint[at];
    } else {
      c = 0.0;
      return dt * Pint[at];
    }
  };
  // difference of the potential between two cells, each part on its own
  auto dif = [](double a, double ac, double b, double bc) -> double {
    if constexpr (SPLIT) return (a - b) + (ac - bc);
    else return a - b;
  };
  double pc;
  double p = pot(ridx(i, j, kb, nxg, nyg), pc);
  bool bad = false;
#pragma unroll
  for (int kk = 0; kk < KC; kk++) {
    const int k = kb + kk;
    if (k > nzg - 1) break;
    if constexpr (SPLIT) {
      const double o = (p + pc) / dt;
      pout[ridx(i, j, k, nxg, nyg)] = o;
      if (i == 2) pout[ridx(nxg - 1, j, k, nxg, nyg)] = o;
      if (k == nzg - 1) {
        pout[ridx(i, j, nzg, nxg, nyg)] = 0.0;
        if (i == 2) pout[ridx(nxg - 1, j, nzg, nxg, nyg)] = 0.0;
      }
    }
    const double pxn = wave_shl1(p);
    const double pxcn = SPLIT ? wave_shl1(pc) : 0.0;
    double pxce;
    const double pxe = pot(ridx(ie + 1, j, k, nxg, nyg), pxce);
    const double px = (i == ie) ? pxe : pxn;
    const double pxc = (i == ie) ? pxce : pxcn;
    const size_t iu = fidx(i, j, k, nx, nyg);
    const double u = U[iu] - dif(px, pxc, p, pc) / hx;
    U[iu] = u;
    bad |= !isfinite(u);
    if (j <= ny - 1) {
      double pyc;
      const double py = pot(ridx(i, j + 1, k, nxg, nyg), pyc);
      const size_t iv = fidx(i, j, k, nxg, ny);
      const double v = V[iv] - dif(py, pyc, p, pc) / hy;
      V[iv] = v;
      bad |= !isfinite(v);
    }
    if (k <= nz - 1) {
      double pzc;
      const double pz = pot(ridx(i, j, k + 1, nxg, nyg), pzc);
      const size_t iw = fidx(i, j, k, nxg, nyg);
      const double w = W[iw] - dif(pz, pzc, p, pc) / g.dzg[k];
      W[iw] = w;
      bad |= !isfinite(w);
      p = pz;
      pc = pzc;
    }
  }
  if (bad && g.nonfinite) atomicOr(g.nonfinite, 1u);
}

// ---------------------------------------------------------------------------
// Boundary conditions, src/boundary_conditions.f90:20-252, one domain.
// ---------------------------------------------------------------------------
// All three boundary-condition passes in one launch.  Applied in the
// reference's order (periodic x, periodic z, walls in y), every boundary
// value ends as a copy of an interior value: the periodic image in x and z
// (x faces 1 -> nx-1, nx -> 2; x centres 1 -> nxg-2, nxg-1 -> 2, nxg -> 3;
// z alike) and, in the y ghost rows of U and W, the negated value of the
// first interior row; V's wall rows are 0.  Each thread writes the boundary
// cells of one (j,k), (i,j) or (i,k) line from those interior values only,
// so no read depends on another write and cells shared by two lines get the
// same value: the result equals the three passes run in order.
__device__ __forceinline__ int img_face(int i, int n) { return i == 1 ? n - 1 : (i == n ? 2 : i); }
__device__ __forceinline__ int img_ctr(int i, int ng) {
  return i == 1 ? ng - 2 : (i == ng - 1 ? 2 : (i == ng ? 3 : i));
}
struct BcMap {
  int nx, ny, nz, nxg, nyg, nzg;
  __device__ double u(const double* U, int i, int j, int k) const {  // U(nx, nyg, nzg)
    const int jj = j == 1 ? 2 : (j == nyg ? nyg - 1 : j);
    const double v = U[fidx(img_face(i, nx), jj, img_ctr(k, nzg), nx, nyg)];
    return jj == j ? v : -v;
  }
  __device__ double v(const double* V, int i, int j, int k) const {  // V(nxg, ny, nzg)
    if (j == 1 || j == ny) return 0.0;
    return V[fidx(img_ctr(i, nxg), j, img_ctr(k, nzg), nxg, ny)];
  }
  __device__ double w(const double* W, int i, int j, int k) const {  // W(nxg, nyg, nz)
    const int jj = j == 1 ? 2 : (j == nyg ? nyg - 1 : j);
    const double v = W[fidx(img_ctr(i, nxg), jj, img_face(k, nz), nxg, nyg)];
    return jj == j ? v : -v;
  }
};

__global__ __launch_bounds__(kBlock) void k_bc_all(DevGrid g, double* __restrict__ U,
                                                   double* __restrict__ V, double* __restrict__ W) {
  const BcMap m{g.nx, g.ny, g.nz, g.nxg, g.nyg, g.nzg};
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg, nzg = g.nzg;
  const long long nX = (long long)nyg * nzg, nZ = (long long)nxg * nyg, nY = (long long)nxg * nzg;
  long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t < nX) {  // x lines: (j, k)
    const int j = (int)(t % nyg) + 1, k = (int)(t / nyg) + 1;
    U[fidx(1, j, k, nx, nyg)] = m.u(U, 1, j, k);
    U[fidx(nx, j, k, nx, nyg)] = m.u(U, nx, j, k);
    if (j <= ny) {
      V[fidx(1, j, k, nxg, ny)] = m.v(V, 1, j, k);
      V[fidx(nxg - 1, j, k, nxg, ny)] = m.v(V, nxg - 1, j, k);
      V[fidx(nxg, j, k, nxg, ny)] = m.v(V, nxg, j, k);
    }
    if (k <= nz) {
      W[fidx(1, j, k, nxg, nyg)] = m.w(W, 1, j, k);
      W[fidx(nxg - 1, j, k, nxg, nyg)] = m.w(W, nxg - 1, j, k);
      W[fidx(nxg, j, k, nxg, nyg)] = m.w(W, nxg, j, k);
    }
    return;
  }
  t -= nX;
  if (t < nZ) {  // z lines: (i, j)
    const int i = (int)(t % nxg) + 1, j = (int)(t / nxg) + 1;
    if (i <= nx) {
      U[fidx(i, j, 1, nx, nyg)] = m.u(U, i, j, 1);
      U[fidx(i, j, nzg - 1, nx, nyg)] = m.u(U, i, j, nzg - 1);
      U[fidx(i, j, nzg, nx, nyg)] = m.u(U, i, j, nzg);
    }
    if (j <= ny) {
      V[fidx(i, j, 1, nxg, ny)] = m.v(V, i, j, 1);
      V[fidx(i, j, nzg - 1, nxg, ny)] = m.v(V, i, j, nzg - 1);
      V[fidx(i, j, nzg, nxg, ny)] = m.v(V, i, j, nzg);
    }
    W[fidx(i, j, 1, nxg, nyg)] = m.w(W, i, j, 1);
    W[fidx(i, j, nz, nxg, nyg)] = m.w(W, i, j, nz);
    return;
  }
  t -= nZ;
  if (t < nY) {  // y lines: (i, k)
    const int i = (int)(t % nxg) + 1, k = (int)(t / nxg) + 1;
    if (i <= nx) {
      U[fidx(i, 1, k, nx, nyg)] = m.u(U, i, 1, k);
      U[fidx(i, nyg, k, nx, nyg)] = m.u(U, i, nyg, k);
    }
    V[fidx(i, 1, k, nxg, ny)] = 0.0;
    V[fidx(i, ny, k, nxg, ny)] = 0.0;
    if (k <= nz) {
      W[fidx(i, 1, k, nxg, nyg)] = m.w(W, i, 1, k);
      W[fidx(i, nyg, k, nxg, nyg)] = m.w(W, i, nyg, k);
    }
  }
}

// ---------------------------------------------------------------------------
// Divergence into rhs_p shape and max|div| reduction (src/monitor.f90:210-221).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double div_at(const DevGrid& g, const double* __restrict__ U,
                                         const double* __restrict__ V,
                                         const double* __restrict__ W, int i, int j, int k) {
  return (U[fidx(i, j, k, g.nx, g.nyg)] - U[fidx(i - 1, j, k, g.nx, g.nyg)]) / g.dxf[i] +
         (V[fidx(i, j, k, g.nxg, g.ny)] - V[fidx(i, j - 1, k, g.nxg, g.ny)]) / g.dyf[j] +
         (W[fidx(i, j, k, g.nxg, g.nyg)] - W[fidx(i, j, k - 1, g.nxg, g.nyg)]) / g.dzf[k];
}

__global__ __launch_bounds__(kBlock) void k_divergence(DevGrid g, const double* __restrict__ U,
                                                       const double* __restrict__ V,
                                                       const double* __restrict__ W,
                                                       double* __restrict__ div) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 2;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 2;
  const int k = blockIdx.z + 2;  // 2..nzg
  if (i > g.nxg - 1 || j > g.nyg - 1) return;
  div[ridx(i, j, k, g.nxg, g.nyg)] = (k <= g.nzg - 1) ? div_at(g, U, V, W, i, j, k) : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_max_div(DevGrid g, const double* __restrict__ U,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ W,
                                                    unsigned long long* __restrict__ red) {
  __shared__ double sm[kBlock / 64];
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);
  const int i = bx * 64 + (threadIdx.x & 63) + 2;
  const int j = by * 4 + (threadIdx.x >> 6) + 2;
  double v = 0.0;
  if (i <= g.nxg - 1 && j <= g.nyg - 1)
    for (int k = bz + 2; k <= g.nzg - 1; k += gridDim.z)
      v = fmax(v, fabs(div_at(g, U, V, W, i, j, k)));
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sm[0];
    for (int w = 1; w < kBlock / 64; w++) m = fmax(m, sm[w]);
    atomicMax(red, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ __launch_bounds__(kBlock) void k_gradient(DevGrid g, const double* __restrict__ p,
                                                     double* __restrict__ Gx,
                                                     double* __restrict__ Gy,
                                                     double* __restrict__ Gz) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 2;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 2;
  const int k = blockIdx.z + 2;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nxg = g.nxg, nyg = g.nyg;
  if (i > nx - 1 || j > nyg - 1) return;
  const double pc = p[ridx(i, j, k, nxg, nyg)];
  Gx[fidx(i, j, k, nx, nyg)] = (p[ridx(i + 1, j, k, nxg, nyg)] - pc) / g.dxg[i];
  if (j <= ny - 1) Gy[fidx(i, j, k, nxg, ny)] = (p[ridx(i, j + 1, k, nxg, nyg)] - pc) / g.dyg[j];
  if (k <= nz - 1) Gz[fidx(i, j, k, nxg, nyg)] = (p[ridx(i, j, k + 1, nxg, nyg)] - pc) / g.dzg[k];
}

// Streaming copy (calibration of bandwidth and traffic counters): one tile
// of 4 x 256 16-byte words per workgroup, non-temporal loads and stores,
// all loads issued first.  Of the shapes measured on a 1 GiB copy (grid-
// stride loops of 2048 to 32768 workgroups, tiles of 4 or 8 words per lane,
// with and without non-temporal hints) this one is fastest.
typedef double vd2 __attribute__((ext_vector_type(2)));
constexpr int kCopyU = 4;
__global__ __launch_bounds__(kBlock) void k_copy(const vd2* __restrict__ src,
                                                 vd2* __restrict__ dst, size_t n) {
  const size_t base = (size_t)blockIdx.x * (kCopyU * kBlock) + threadIdx.x;
  vd2 v[kCopyU];
#pragma unroll
  for (int u = 0; u < kCopyU; u++) {
    const size_t q = base + (size_t)u * kBlock;
    if (q < n) v[u] = __builtin_nontemporal_load(src + q);
  }
#pragma unroll
  for (int u = 0; u < kCopyU; u++) {
    const size_t q = base + (size_t)u * kBlock;
    if (q < n) __builtin_nontemporal_store(v[u], dst + q);
  }
}

// Dispatch a compile-time FFT length.
#define C3D_DISPATCH_N(n, BODY)                     \
  switch (n) {                                      \
    case 16: { constexpr int NN = 16; BODY; } break;   \
    case 32: { constexpr int NN = 32; BODY; } break;   \
    case 48: { constexpr int NN = 48; BODY; } break;   \
    case 64: { constexpr int NN = 64; BODY; } break;   \
    case 96: { constexpr int NN = 96; BODY; } break;   \
    case 128: { constexpr int NN = 128; BODY; } break; \
    case 192: { constexpr int NN = 192; BODY; } break; \
    case 256: { constexpr int NN = 256; BODY; } break; \
    case 384: { constexpr int NN = 384; BODY; } break; \
    case 512: { constexpr int NN = 512; BODY; } break; \
    default: return hipErrorInvalidValue;           \
  }

#define C3D_DISPATCH_TK(tk, BODY)                  \
  switch (tk) {                                    \
    case 16: { constexpr int TK = 16; BODY; } break; \
    case 8: { constexpr int TK = 8; BODY; } break;   \
    case 4: { constexpr int TK = 4; BODY; } break;   \
    case 2: { constexpr int TK = 2; BODY; } break;   \
    default: return hipErrorInvalidValue;          \
  }

inline dim3 grid_pts(const DevGrid& g, int ni, int nj, int nk) {
  return dim3((ni + 63) / 64, (nj + 3) / 4, nk);
}

}  // namespace

hipError_t launch_fwd_x(const DevGrid& g, bool from_uvw, const double* U, const double* V,
                        const double* W, const double* rhs, cpx<double>* S, hipStream_t s) {
  const int tk = tile_rows(g.nxp, g.nzp);
  dim3 grid(g.nzp / tk, g.nym);
  // U, V, W together below 128 MB stay in the 256 MB last-level cache
  // between the kernel that wrote them and this one: cached loads there
  // (measured 13.2 against 15.6 us at 128^3), non-temporal loads above.
  const bool small = 3.0 * g.nx * g.nyg * g.nzg * 8.0 < 128.0 * (1 << 20);
  C3D_DISPATCH_N(g.nxp, C3D_DISPATCH_TK(tk, {
    if (from_uvw && small)
      hipLaunchKernelGGL((k_fwd_x<double, NN, TK, kSrcUVWc>), grid, dim3(kBlock), 0, s, g, U, V, W,
                         (const void*)rhs, g.twx64, S, nullptr, nullptr, nullptr);
    else if (from_uvw)
      hipLaunchKernelGGL((k_fwd_x<double, NN, TK, kSrcUVW>), grid, dim3(kBlock), 0, s, g, U, V, W,
                         (const void*)rhs, g.twx64, S, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_fwd_x<double, NN, TK, kSrcRhs>), grid, dim3(kBlock), 0, s, g, U, V, W,
                         (const void*)rhs, g.twx64, S, nullptr, nullptr, nullptr);
  }));
  return hipGetLastError();
}

hipError_t launch_fwd_x32(const DevGrid& g, const double* U, const double* V, const double* W,
                          double* b, cpx<float>* S, hipStream_t s) {
  if (!U || !V || !W) return hipErrorInvalidValue;
  const int tk = tile_rows(g.nxp, g.nzp);
  dim3 grid(g.nzp / tk, g.nym);
  C3D_DISPATCH_N(g.nxp, C3D_DISPATCH_TK(tk, {
    if (b)
      hipLaunchKernelGGL((k_fwd_x<float, NN, TK, kSrcUVWB>), grid, dim3(kBlock), 0, s, g, U, V, W,
                         (const void*)nullptr, g.twx32, S, b, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_fwd_x<float, NN, TK, kSrcUVW>), grid, dim3(kBlock), 0, s, g, U, V, W,
                         (const void*)nullptr, g.twx32, S, nullptr, nullptr, nullptr);
  }));
  return hipGetLastError();
}

namespace {
template <typename real>
hipError_t fwd_x_resid(const DevGrid& g, const double* b, const double* phi, double* gauge,
                       const cpx<real>* tw, cpx<real>* S, unsigned long long* rmax, double* rout,
                       hipStream_t s) {
  const int tk = tile_rows(g.nxp, g.nzp);
  dim3 grid(g.nzp / tk, g.nym);
  // gauge[0] = result, gauge[1 .. nzp] = row sums
  hipLaunchKernelGGL(k_gauge_rows, dim3(g.nzp), dim3(kBlock), 0, s, g, phi, gauge + 1);
  hipLaunchKernelGGL(k_gauge_final, dim3(1), dim3(kBlock), 0, s, g, (const double*)(gauge + 1), gauge);
  C3D_DISPATCH_N(g.nxp, C3D_DISPATCH_TK(tk, {
    hipLaunchKernelGGL((k_fwd_x<real, NN, TK, kSrcResid>), grid, dim3(kBlock), 0, s, g, phi,
                       nullptr, nullptr, (const void*)b, tw, S, rout, (const double*)gauge, rmax);
  }));
  return hipGetLastError();
}
}  // namespace

hipError_t launch_fwd_x_resid32(const DevGrid& g, const double* b, const double* phi, double* gauge,
                                cpx<float>* S, unsigned long long* rmax, double* rout, hipStream_t s) {
  return fwd_x_resid<float>(g, b, phi, gauge, g.twx32, S, rmax, rout, s);
}

hipError_t launch_fwd_x_resid64(const DevGrid& g, const double* b, const double* phi, double* gauge,
                                cpx<double>* S, unsigned long long* rmax, double* rout,
                                hipStream_t s) {
  return fwd_x_resid<double>(g, b, phi, gauge, g.twx64, S, rmax, rout, s);
}

hipError_t launch_fwd_x_b64(const DevGrid& g, const double* U, const double* V, const double* W,
                            double* b, cpx<double>* S, hipStream_t s) {
  const int tk = tile_rows(g.nxp, g.nzp);
  dim3 grid(g.nzp / tk, g.nym);
  C3D_DISPATCH_N(g.nxp, C3D_DISPATCH_TK(tk, {
    hipLaunchKernelGGL((k_fwd_x<double, NN, TK, kSrcUVWB>), grid, dim3(kBlock), 0, s, g, U, V, W,
                       (const void*)nullptr, g.twx64, S, b, nullptr, nullptr);
  }));
  return hipGetLastError();
}

// Threads per column of the z transforms: 64 (4 columns per workgroup).
// C3D_ZNT32=32 gives the FP32 transforms 32 threads and 8 columns per
// workgroup instead (tuning knob).
static int fft_z_nt32() {
  static const int nt = [] {
    const char* e = getenv("C3D_ZNT32");
    const int v = e ? atoi(e) : 0;
    return (v == 32 || v == 64) ? v : 64;
  }();
  return nt;
}

template <typename real>
static hipError_t fft_z_any(const DevGrid& g, int sign, cpx<real>* S, const cpx<real>* tw,
                            hipStream_t s) {
  const int ncol = g.nym * g.nxh;
  const int nt = (sizeof(real) == 4) ? fft_z_nt32() : 64;
  const int nc = kBlock / nt;
  dim3 grid((ncol + nc - 1) / nc);
  C3D_DISPATCH_N(g.nzp, {
    if (nt == 32 && sign < 0)
      hipLaunchKernelGGL((k_fft_z<real, NN, -1, 32>), grid, dim3(kBlock), 0, s, S, tw, ncol);
    else if (nt == 32)
      hipLaunchKernelGGL((k_fft_z<real, NN, +1, 32>), grid, dim3(kBlock), 0, s, S, tw, ncol);
    else if (sign < 0)
      hipLaunchKernelGGL((k_fft_z<real, NN, -1, 64>), grid, dim3(kBlock), 0, s, S, tw, ncol);
    else
      hipLaunchKernelGGL((k_fft_z<real, NN, +1, 64>), grid, dim3(kBlock), 0, s, S, tw, ncol);
  });
  return hipGetLastError();
}

hipError_t launch_fft_z(const DevGrid& g, int sign, cpx<double>* S, hipStream_t s) {
  return fft_z_any<double>(g, sign, S, g.twz64, s);
}
hipError_t launch_fft_z32(const DevGrid& g, int sign, cpx<float>* S, hipStream_t s) {
  return fft_z_any<float>(g, sign, S, g.twz32, s);
}

#define C3D_DISPATCH_M(nym, BODY)                    \
  switch (nym) {                                     \
    case 64: { constexpr int MM = 8; BODY; } break;    \
    case 128: { constexpr int MM = 16; BODY; } break;  \
    case 192: { constexpr int MM = 24; BODY; } break;  \
    default: return hipErrorInvalidValue;            \
  }

// 16 parts of nym/16 rows (152 VGPRs at nym = 192: 3 waves per SIMD against
// 2 with 8 parts of 24 rows).
#define C3D_DISPATCH_M16(nym, BODY)                  \
  switch (nym) {                                     \
    case 64: { constexpr int MM = 4; BODY; } break;    \
    case 128: { constexpr int MM = 8; BODY; } break;   \
    case 192: { constexpr int MM = 12; BODY; } break;  \
    default: return hipErrorInvalidValue;            \
  }

hipError_t launch_tri(const DevGrid& g, cpx<double>* S, double* CP, hipStream_t s) {
  const size_t nm = (size_t)g.nxh * g.nzp;
  const dim3 grid((unsigned)((nm + kBlock - 1) / kBlock));
  const dim3 gridp((unsigned)((nm + kBlock / 8 - 1) / (kBlock / 8)));
  if (!tri_part_supported(g)) {
    hipLaunchKernelGGL(k_tri64, grid, dim3(kBlock), 0, s, g, S, CP);
    return hipGetLastError();
  }
  C3D_DISPATCH_M(g.nym, hipLaunchKernelGGL((k_tri_part<double, double, MM, kTriAll>), gridp,
                                           dim3(kBlock), 0, s, g, S, nullptr, 0, 0.0, S));
  return hipGetLastError();
}

bool tri_part_supported(const DevGrid& g) { return g.nym == 64 || g.nym == 128 || g.nym == 192; }

namespace {
// The FP32 solves of launch_tri32 in P parts of MM rows (16 modes per wave
// when P = 16).
template <int P, int MM>
hipError_t tri32_parts(const DevGrid& g, cpx<float>* S, const int* low, int nlow, double lam_thr,
                       hipStream_t s) {
  constexpr int MWX = P == 16 ? 16 : 0;
  constexpr int MB = kBlock / P;
  const size_t nm = (size_t)g.nxh * g.nzp;
  const dim3 grid((unsigned)((nm + MB - 1) / MB));
  if (!low) {  // FP64 arithmetic on every mode
    hipLaunchKernelGGL((k_tri_part<float, double, MM, kTriAll, float, P, MWX>), grid, dim3(kBlock),
                       0, s, g, S, nullptr, 0, 0.0, S);
    return hipGetLastError();
  }
  if (nlow <= 0) return hipErrorInvalidValue;  // the (0,0) mode is always a low mode
  const dim3 gridl((unsigned)((nlow + MB - 1) / MB));
  hipLaunchKernelGGL((k_tri_part<float, float, MM, kTriHigh, float, P, MWX>), grid, dim3(kBlock), 0,
                     s, g, S, nullptr, 0, lam_thr, S);
  hipLaunchKernelGGL((k_tri_part<float, double, MM, kTriList, float, P, MWX>), gridl, dim3(kBlock),
                     0, s, g, S, low, nlow, 0.0, S);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_tri32(const DevGrid& g, cpx<float>* S, const int* low, int nlow, double lam_thr,
                        hipStream_t s) {
  // 192 rows as 16 parts of 12: 8 parts of 24 with all of a lane's FP32 rows
  // in flight need more than 256 registers (one wave per SIMD)
  switch (g.nym) {
    case 64: return tri32_parts<8, 8>(g, S, low, nlow, lam_thr, s);
    case 128: return tri32_parts<8, 16>(g, S, low, nlow, lam_thr, s);
    case 192: return tri32_parts<16, 12>(g, S, low, nlow, lam_thr, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_tri32to64(const DevGrid& g, cpx<float>* S32, cpx<double>* S64, hipStream_t s) {
  // FP32 right-hand side, FP64 solution: 16 parts of nym/16 rows, 16 modes
  // per wave (94 against 99 us with 8 parts at 384x193x384; the FP64 solve
  // gains nothing from it and keeps 8 parts)
  const size_t nm = (size_t)g.nxh * g.nzp;
  constexpr int MB = kBlock / 16;
  const dim3 grid((unsigned)((nm + MB - 1) / MB));
  C3D_DISPATCH_M16(g.nym, hipLaunchKernelGGL((k_tri_part<float, double, MM, kTriAll, double, 16, 16>),
                                             grid, dim3(kBlock), 0, s, g, S32, nullptr, 0, 0.0, S64));
  return hipGetLastError();
}

hipError_t launch_inv_x(const DevGrid& g, int mode, const cpx<double>* S, double* out, double* P,
                        double dt, bool rk1, hipStream_t s) {
  if (mode != 0 && mode != 1) return hipErrorInvalidValue;
  return launch_inv_x64_ref(g, mode, S, nullptr, nullptr, out, nullptr, P, dt, rk1, s);
}

namespace {
#define C3D_INV_CASE(MD)                                                                          \
  case MD:                                                                                        \
    hipLaunchKernelGGL((k_inv_x<real, NN, TK, MD>), grid, dim3(kBlock), 0, s, g, S, tw, out, P, dt, \
                       rk1 ? 1 : 0, phi_in, phi_in2, out2);                                       \
    break;
template <typename real>
hipError_t inv_x_modes(const DevGrid& g, int mode, const cpx<real>* S, const cpx<real>* tw,
                       const double* phi_in, const double* phi_in2, double* out, double* out2,
                       double* P, double dt, bool rk1, hipStream_t s) {
  if ((mode == 3 && !phi_in) || (mode == 5 && (!out2 || (rk1 && !phi_in))))
    return hipErrorInvalidValue;
  const int tk = tile_rows(g.nxp, g.nzp);
  dim3 grid(g.nzp / tk, g.nym);
  C3D_DISPATCH_N(g.nxp, C3D_DISPATCH_TK(tk, {
    switch (mode) {
      C3D_INV_CASE(0)
      C3D_INV_CASE(1)
      C3D_INV_CASE(3)
      C3D_INV_CASE(5)
      default: return hipErrorInvalidValue;
    }
  }));
  return hipGetLastError();
}
#undef C3D_INV_CASE
}  // namespace

hipError_t launch_inv_x64_ref(const DevGrid& g, int mode, const cpx<double>* S, const double* phi_in,
                              const double* phi_in2, double* out, double* out2, double* P, double dt,
                              bool rk1, hipStream_t s) {
  return inv_x_modes<double>(g, mode, S, g.twx64, phi_in, phi_in2, out, out2, P, dt, rk1, s);
}

hipError_t launch_inv_x32(const DevGrid& g, int mode, const cpx<float>* S, const double* phi_in,
                          const double* phi_in2, double* out, double* out2, double* P, double dt,
                          bool rk1, hipStream_t s) {
  return inv_x_modes<float>(g, mode, S, g.twx32, phi_in, phi_in2, out, out2, P, dt, rk1, s);
}

hipError_t launch_p_ghosts(const DevGrid& g, double* P, hipStream_t s) {
  hipLaunchKernelGGL(k_pghost_x, dim3((g.nyg + 63) / 64, g.nzg), dim3(64), 0, s, g, P);
  hipLaunchKernelGGL(k_pghost_y, dim3((g.nxg + 63) / 64, g.nzg), dim3(64), 0, s, g, P);
  hipLaunchKernelGGL(k_pghost_z, dim3((g.nxg + 63) / 64, g.nyg), dim3(64), 0, s, g, P);
  return hipGetLastError();
}

// z planes per lane of the gradient correction (at 384x193x384, 2 measured
// fastest for the single potential: 0.303 against 0.305 ms for 1, more
// workgroups in flight; 4 for the two-part potential: 0.457 against
// 0.468 ms for 2); C3D_GRAD_KC (1, 2, 4, 8, 16) overrides both (tuning knob).
static int grad_kc(bool split) {
  static const int forced = [] {
    const char* e = getenv("C3D_GRAD_KC");
    const int v = e ? atoi(e) : 0;
    return (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) ? v : 0;
  }();
  return forced ? forced : (split ? kGradKCSplitDefault : kGradKCDefault);
}

namespace {
template <bool SPLIT>
hipError_t grad_correct_any(const DevGrid& g, double* U, double* V, double* W, const double* Pint,
                            const double* phic, double* pout, double dt, hipStream_t s) {
  const int kc = grad_kc(SPLIT);
  const dim3 grid = grid_pts(g, g.nx - 2, g.nyg - 2, (g.nzg - 2 + kc - 1) / kc);
  switch (kc) {
#define C3D_GC_CASE(KC)                                                                         \
  case KC:                                                                                      \
    hipLaunchKernelGGL((k_grad_correct<KC, SPLIT>), grid, dim3(kBlock), 0, s, g, U, V, W, Pint,     \
                       phic, pout, dt);                                                         \
    break;
    C3D_GC_CASE(1)
    C3D_GC_CASE(2)
    C3D_GC_CASE(4)
    C3D_GC_CASE(8)
    C3D_GC_CASE(16)
#undef C3D_GC_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_grad_correct(const DevGrid& g, double* U, double* V, double* W, const double* Pint,
                               double dt, hipStream_t s) {
  return grad_correct_any<false>(g, U, V, W, Pint, nullptr, nullptr, dt, s);
}

hipError_t launch_grad_correct_split(const DevGrid& g, double* U, double* V, double* W,
                                     const double* phi1, const double* phic, double* Pint, double dt,
                                     hipStream_t s) {
  return grad_correct_any<true>(g, U, V, W, phi1, phic, Pint, dt, s);
}

hipError_t launch_apply_bc(const DevGrid& g, double* U, double* V, double* W, hipStream_t s) {
  const long long n = (long long)g.nyg * g.nzg + (long long)g.nxg * g.nyg + (long long)g.nxg * g.nzg;
  hipLaunchKernelGGL(k_bc_all, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g, U,
                     V, W);
  return hipGetLastError();
}


hipError_t launch_divergence(const DevGrid& g, const double* U, const double* V, const double* W,
                             double* div, hipStream_t s) {
  hipLaunchKernelGGL(k_divergence, grid_pts(g, g.nxg - 2, g.nyg - 2, g.nzg - 1), dim3(kBlock), 0,
                     s, g, U, V, W, div);
  return hipGetLastError();
}

hipError_t launch_gradient(const DevGrid& g, const double* p, double* Gx, double* Gy, double* Gz,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_gradient, grid_pts(g, g.nx - 2, g.nyg - 2, g.nzg - 2), dim3(kBlock), 0, s,
                     g, p, Gx, Gy, Gz);
  return hipGetLastError();
}

hipError_t launch_max_div(const DevGrid& g, const double* U, const double* V, const double* W,
                          unsigned long long* red, hipStream_t s) {
  hipLaunchKernelGGL(k_max_div, grid_pts(g, g.nxg - 2, g.nyg - 2, g.nzg - 2 < 16 ? g.nzg - 2 : 16),
                     dim3(kBlock), 0, s, g, U, V, W, red);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t nbytes, hipStream_t s) {
  const size_t n = nbytes / sizeof(vd2);
  const size_t tiles = (n + kCopyU * kBlock - 1) / (kCopyU * kBlock);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)tiles), dim3(kBlock), 0, s, (const vd2*)src, (vd2*)dst, n);
  return hipGetLastError();
}

}  // namespace c3d
