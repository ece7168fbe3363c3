// Stockham autosort FFT on LDS-resident rows, for wave64 workgroups on gfx950.
//
// A length-N complex row lives in LDS at padded positions lp(i) (see below);
// NT threads of the workgroup own it.
// Each radix-R pass loads its R inputs per butterfly into registers, syncs,
// applies the twiddles and the radix-R DFT and stores the autosorted
// outputs back into the same LDS row (so a row needs N complex of LDS, not
// 2N).  Twiddles come from a per-length table tw[m] = exp(-2 pi i m / N)
// (rounded once from extended precision on the host); the inverse
// transform uses conj(tw).  Unnormalised, natural order in and out:
//   forward  (SIGN = -1): X_k = sum_n x_n exp(-2 pi i n k / N)
//   backward (SIGN = +1): x_n = sum_k X_k exp(+2 pi i n k / N)
// which is FFTW's FFTW_FORWARD / FFTW_BACKWARD convention.
//
// Every pass ends in __syncthreads(), so all threads of the workgroup must
// call the transform together (rows may be owned by different thread groups).
#pragma once

#include "c3d_common.hpp"

namespace c3d {

// Padded LDS position of element i of a row: one spare complex after every
// 8, so the stride-R stores of a first radix-8 pass (and stride-8 patterns
// in general) spread over all LDS banks.  A padded row occupies
// lds_row_len(N) complex.
__host__ __device__ __forceinline__ constexpr int lp(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int lds_row_len(int n) { return n + n / 8; }

template <typename real, int R, int SIGN>
__device__ __forceinline__ void dft_small(cpx<real>* v) {
  if constexpr (R == 2) {
    cpx<real> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 3) {
    const real s3 = (real)0.86602540378443864676372317075294;  // sqrt(3)/2
    cpx<real> t = v[1] + v[2], d = v[1] - v[2];
    cpx<real> m = cmk(v[0].x - (real)0.5 * t.x, v[0].y - (real)0.5 * t.y);
    cpx<real> r = cmk((real)(-SIGN) * s3 * d.y, (real)SIGN * s3 * d.x);  // SIGN*i*s3*d
    v[0] = v[0] + t;
    v[1] = m + r;
    v[2] = m - r;
  } else if constexpr (R == 4) {
    cpx<real> a0 = v[0] + v[2], a1 = v[0] - v[2];
    cpx<real> a2 = v[1] + v[3], d = v[1] - v[3];
    cpx<real> a3 = cmk((real)(-SIGN) * d.y, (real)SIGN * d.x);  // SIGN*i*d
    v[0] = a0 + a2;
    v[2] = a0 - a2;
    v[1] = a1 + a3;
    v[3] = a1 - a3;
  } else if constexpr (R == 6) {
    // even and odd samples through two 3-point DFTs, then
    // X_k = A_k + w6^k B_k, X_{k+3} = A_k - w6^k B_k (k = 0, 1, 2) with
    // w6 = exp(SIGN 2 pi i / 6): w6^1 = 1/2 + SIGN i sqrt(3)/2,
    // w6^2 = -1/2 + SIGN i sqrt(3)/2
    const real s3 = (real)0.86602540378443864676372317075294;  // sqrt(3)/2
    cpx<real> a[3] = {v[0], v[2], v[4]};
    cpx<real> b[3] = {v[1], v[3], v[5]};
    dft_small<real, 3, SIGN>(a);
    dft_small<real, 3, SIGN>(b);
    const cpx<real> b1 = cmk((real)0.5 * b[1].x - (real)SIGN * s3 * b[1].y,
                             (real)0.5 * b[1].y + (real)SIGN * s3 * b[1].x);
    const cpx<real> b2 = cmk((real)-0.5 * b[2].x - (real)SIGN * s3 * b[2].y,
                             (real)-0.5 * b[2].y + (real)SIGN * s3 * b[2].x);
    v[0] = a[0] + b[0];
    v[3] = a[0] - b[0];
    v[1] = a[1] + b1;
    v[4] = a[1] - b1;
    v[2] = a[2] + b2;
    v[5] = a[2] - b2;
  } else if constexpr (R == 8) {
    const real c = (real)0.70710678118654752440084436210485;  // sqrt(2)/2
    cpx<real> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<real> o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<real, 4, SIGN>(e);
    dft_small<real, 4, SIGN>(o);
    // w8^m = exp(SIGN 2 pi i m / 8)
    cpx<real> o1 = cmk(c * (o[1].x - (real)SIGN * o[1].y), c * (o[1].y + (real)SIGN * o[1].x));
    cpx<real> o2 = cmk((real)(-SIGN) * o[2].y, (real)SIGN * o[2].x);
    cpx<real> o3 = cmk(c * (-o[3].x - (real)SIGN * o[3].y), c * (-o[3].y + (real)SIGN * o[3].x));
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
}

// One radix-R pass; NS = product of the radices already applied.
template <typename real, int N, int NT, int R, int NS, int SIGN>
__device__ __forceinline__ void stockham_pass(cpx<real>* buf, const cpx<real>* __restrict__ tw,
                                              int t) {
  constexpr int NB = N / R;
  constexpr int PER = (NB + NT - 1) / NT;
  cpx<real> v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; b++) {
    const int j = t + b * NT;
    if (PER * NT == NB || j < NB) {
#pragma unroll
      for (int r = 0; r < R; r++) v[b][r] = buf[lp(j + r * NB)];
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; r++) {
          const cpx<real> w = tw[r * k * (N / (NS * R))];
          v[b][r] = (SIGN < 0) ? cmul(v[b][r], w) : cmulc(v[b][r], w);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; b++) {
    const int j = t + b * NT;
    if (PER * NT == NB || j < NB) {
      dft_small<real, R, SIGN>(v[b]);
      const int k = j % NS;
      const int idx = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[lp(idx + r * NS)] = v[b][r];
    }
  }
  __syncthreads();
}

template <typename real, int N, int NT, int SIGN, int NS, int... Rs>
struct fft_passes;

template <typename real, int N, int NT, int SIGN, int NS>
struct fft_passes<real, N, NT, SIGN, NS> {
  __device__ static __forceinline__ void run(cpx<real>*, const cpx<real>*, int) {}
};

template <typename real, int N, int NT, int SIGN, int NS, int R, int... Rs>
struct fft_passes<real, N, NT, SIGN, NS, R, Rs...> {
  __device__ static __forceinline__ void run(cpx<real>* buf, const cpx<real>* tw, int t) {
    stockham_pass<real, N, NT, R, NS, SIGN>(buf, tw, t);
    fft_passes<real, N, NT, SIGN, NS * R, Rs...>::run(buf, tw, t);
  }
};

// Lengths supported by the library (grid sizes nxp, nzp).
__host__ __device__ constexpr bool fft_length_supported(int n) {
  return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 192 ||
         n == 256 || n == 384 || n == 512;
}

template <typename real, int N, int NT, int SIGN>
__device__ __forceinline__ void fft_lds(cpx<real>* buf, const cpx<real>* __restrict__ tw, int t) {
  if constexpr (N == 16) fft_passes<real, N, NT, SIGN, 1, 4, 4>::run(buf, tw, t);
  else if constexpr (N == 32) fft_passes<real, N, NT, SIGN, 1, 8, 4>::run(buf, tw, t);
  else if constexpr (N == 48) fft_passes<real, N, NT, SIGN, 1, 4, 4, 3>::run(buf, tw, t);
  else if constexpr (N == 64) fft_passes<real, N, NT, SIGN, 1, 8, 8>::run(buf, tw, t);
  else if constexpr (N == 96) fft_passes<real, N, NT, SIGN, 1, 8, 4, 3>::run(buf, tw, t);
  else if constexpr (N == 128) fft_passes<real, N, NT, SIGN, 1, 8, 4, 4>::run(buf, tw, t);
  else if constexpr (N == 192) fft_passes<real, N, NT, SIGN, 1, 8, 8, 3>::run(buf, tw, t);
  else if constexpr (N == 256) fft_passes<real, N, NT, SIGN, 1, 8, 8, 4>::run(buf, tw, t);
  else if constexpr (N == 384) fft_passes<real, N, NT, SIGN, 1, 8, 8, 6>::run(buf, tw, t);
  else if constexpr (N == 512) fft_passes<real, N, NT, SIGN, 1, 8, 8, 8>::run(buf, tw, t);
  else static_assert(N == 16, "unsupported FFT length");
}

}  // namespace c3d
