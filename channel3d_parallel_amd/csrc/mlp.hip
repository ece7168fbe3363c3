// Learned initial guess for the IB Schur solve: a 54 -> 256 -> 256 -> 256 -> 4
// ReLU MLP with hand-written training and inference kernels for gfx950.
//
// GEMMs run on the matrix cores: v_mfma_f32_16x16x16_bf16 with BF16 operands
// (round-to-nearest-even from the FP32 activations and master weights) and
// FP32 accumulation, on 64 x 64 output tiles staged through LDS.
//
// MFMA 16x16x16 operand layout (lane l, r = l % 16, q = l / 16):
//   A: 4 values A(row r, k = 4q .. 4q+3)      B: 4 values B(k = 4q .. 4q+3, col r)
//   D: 4 values D(row 4q .. 4q+3, col r)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace c3d {
namespace mlp {

namespace {

constexpr int kBlock = 256;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// C(M x N) = op(A)(M x K) * op(B)(K x N), FP32 in memory, BF16 MFMA.
//   op(A)(m,k) = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// Epilogues: kEpiNone, kEpiBias (+bias[n]), kEpiBiasRelu, kEpiMask
// (C *= (H(m,n) > 0), the ReLU derivative of the forward activation H).
//
// A workgroup computes a T x T tile of C (T = 64, or 32 for products too
// small to give every CU a 64 x 64 tile); its 4 wavefronts each own a
// T/2 x T/2 quarter as (T/32) x (T/32) MFMA tiles.  K advances in slices of
// 32: the workgroup converts the T x 32 slices of op(A) and op(B) to BF16
// into LDS (row-major [m][k] and [n][k], so both operand fragments are 4
// consecutive k of one row), and every wavefront then reads its fragments
// from LDS.  The next slice's loads are issued into registers before the
// current slice's MFMAs; with VA / VB the operand is read as float4 along
// its contiguous index (that extent and the leading dimension multiples of
// 4, the base 16-byte aligned).
constexpr int kTK = 32;
constexpr int kTKP = kTK + 4;  // LDS row pitch in BF16 elements (72 bytes)

// Slice of an operand X with rows r (m or n) and reductions k: KC means
// X[r*ld + k] (k contiguous), else X[k*ld + r].  E values per thread.
template <bool KC, bool V, int T>
struct Slice {
  static constexpr int E = T * kTK / kBlock;
  float v[E];
  __device__ __forceinline__ void load(const float* __restrict__ X, int ld, int R, int K, int r0,
                                       int k0) {
    const int tid = threadIdx.x;
    if constexpr (V) {
#pragma unroll
      for (int g = 0; g < E / 4; g++) {
        const int e = tid + g * kBlock;  // float4 group
        int rr, kk;
        if constexpr (KC) { rr = e / (kTK / 4); kk = 4 * (e % (kTK / 4)); }
        else { kk = e / (T / 4); rr = 4 * (e % (T / 4)); }
        const int r = r0 + rr, k = k0 + kk;
        float4_t x = float4_t{0.f, 0.f, 0.f, 0.f};
        if (r < R && k < K)
          x = *reinterpret_cast<const float4_t*>(KC ? &X[(size_t)r * ld + k] : &X[(size_t)k * ld + r]);
#pragma unroll
        for (int c = 0; c < 4; c++) v[4 * g + c] = x[c];
      }
    } else {
#pragma unroll
      for (int i = 0; i < E; i++) {
        const int e = tid + i * kBlock;
        const int rr = KC ? e / kTK : e % T, kk = KC ? e % kTK : e / T;
        const int r = r0 + rr, k = k0 + kk;
        v[i] = (r < R && k < K) ? (KC ? X[(size_t)r * ld + k] : X[(size_t)k * ld + r]) : 0.f;
      }
    }
  }
  __device__ __forceinline__ void store(short* Xs) const {
    const int tid = threadIdx.x;
    if constexpr (V) {
#pragma unroll
      for (int g = 0; g < E / 4; g++) {
        const int e = tid + g * kBlock;
        if constexpr (KC) {
          const int rr = e / (kTK / 4), kk = 4 * (e % (kTK / 4));
          *reinterpret_cast<short4_t*>(&Xs[rr * kTKP + kk]) =
              short4_t{bf16_bits(v[4 * g]), bf16_bits(v[4 * g + 1]), bf16_bits(v[4 * g + 2]),
                       bf16_bits(v[4 * g + 3])};
        } else {
          const int kk = e / (T / 4), rr = 4 * (e % (T / 4));
#pragma unroll
          for (int c = 0; c < 4; c++) Xs[(rr + c) * kTKP + kk] = bf16_bits(v[4 * g + c]);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < E; i++) {
        const int e = tid + i * kBlock;
        const int rr = KC ? e / kTK : e % T, kk = KC ? e % kTK : e / T;
        Xs[rr * kTKP + kk] = bf16_bits(v[i]);
      }
    }
  }
};

template <bool TA, bool TB, int EPI, int T, bool VA, bool VB>
__global__ __launch_bounds__(kBlock) void k_gemm(int M, int N, int K, const float* __restrict__ A,
                                                 int lda, const float* __restrict__ B, int ldb,
                                                 float* __restrict__ C, int ldc,
                                                 const float* __restrict__ bias,
                                                 const float* __restrict__ H, int ldh) {
  constexpr int F = T / 32;  // MFMA tiles per wavefront along m and along n
  __shared__ __align__(16) short As[T * kTKP];
  __shared__ __align__(16) short Bs[T * kTKP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.x * T, n0 = blockIdx.y * T;
  const int wm = (wave >> 1) * (T / 2), wn = (wave & 1) * (T / 2);
  const int r = lane & 15, q = lane >> 4;
  float4_t acc[F][F];
#pragma unroll
  for (int i = 0; i < F; i++)
#pragma unroll
    for (int j = 0; j < F; j++) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  Slice<!TA, VA, T> sa;
  Slice<TB, VB, T> sb;
  sa.load(A, lda, M, K, m0, 0);
  sb.load(B, ldb, N, K, n0, 0);
  for (int k0 = 0; k0 < K; k0 += kTK) {
    sa.store(As);
    sb.store(Bs);
    __syncthreads();
    if (k0 + kTK < K) {  // next slice in flight during this slice's MFMAs
      sa.load(A, lda, M, K, m0, k0 + kTK);
      sb.load(B, ldb, N, K, n0, k0 + kTK);
    }
#pragma unroll
    for (int ks = 0; ks < kTK; ks += 16) {
      short4_t a[F], b[F];
#pragma unroll
      for (int i = 0; i < F; i++)
        a[i] = *reinterpret_cast<const short4_t*>(&As[(wm + 16 * i + r) * kTKP + ks + 4 * q]);
#pragma unroll
      for (int j = 0; j < F; j++)
        b[j] = *reinterpret_cast<const short4_t*>(&Bs[(wn + 16 * j + r) * kTKP + ks + 4 * q]);
#pragma unroll
      for (int i = 0; i < F; i++)
#pragma unroll
        for (int j = 0; j < F; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();  // the slices are overwritten next
  }
#pragma unroll
  for (int i = 0; i < F; i++)
#pragma unroll
    for (int j = 0; j < F; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int m = m0 + wm + 16 * i + 4 * q + e, n = n0 + wn + 16 * j + r;
        if (m < M && n < N) {
          float v = acc[i][j][e];
          if (EPI == kEpiBias || EPI == kEpiBiasRelu) v += bias[n];
          if (EPI == kEpiBiasRelu) v = fmaxf(v, 0.f);
          if (EPI == kEpiMask) v = (H[(size_t)m * ldh + n] > 0.f) ? v : 0.f;
          C[(size_t)m * ldc + n] = v;
        }
      }
}

// Mean-squared error over n x 4 outputs: dY = 2 (Y - T) / (4n); per-block
// partial sums of (Y - T)^2 / (4n) in part.
__global__ __launch_bounds__(kBlock) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                                int count, float* __restrict__ dY,
                                                float* __restrict__ part) {
  __shared__ float sh[kBlock];
  const float inv = 1.f / (float)count;
  float acc = 0.f;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < count; e += gridDim.x * kBlock) {
    const float d = Y[e] - T[e];
    dY[e] = 2.f * d * inv;
    acc += d * d * inv;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// db[n] = sum over rows of dY(m, n), fixed order: a workgroup owns
// kColN columns; its kColR row slices each sum rows m = slice, slice + kColR,
// ... (consecutive lanes read consecutive columns), then the slices are
// added pairwise in LDS.
constexpr int kColN = 16, kColR = kBlock / kColN;
__global__ __launch_bounds__(kBlock) void k_colsum(const float* __restrict__ dY, int M, int N,
                                                   float* __restrict__ db) {
  __shared__ float sh[kColR][kColN];
  const int c = threadIdx.x % kColN, r = threadIdx.x / kColN;
  const int n = blockIdx.x * kColN + c;
  float s = 0.f;
  if (n < N)
    for (int m = r; m < M; m += kColR) s += dY[(size_t)m * N + n];
  sh[r][c] = s;
  __syncthreads();
  for (int h = kColR / 2; h > 0; h >>= 1) {
    if (r < h) sh[r][c] += sh[r + h][c];
    __syncthreads();
  }
  if (r == 0 && n < N) db[n] = sh[0][c];
}

// Adam (FP32), bias-corrected step size lr_t passed in.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr_t, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// Feature rows of the IB markers: for neighbour slot s (0..8) of marker l,
// channels (r_u x, r_u y, r_u z, r_p, dudn_x, dudn_y) of marker nb9[9l+s] --
// the residual b - S fprev of the warm start and the previous force --
// divided by the channel scale.
__global__ void k_features(int nb, const int* __restrict__ nb9, const double* __restrict__ res,
                           const double* __restrict__ fprev, const double* __restrict__ nrm,
                           const float* __restrict__ iscale, float* __restrict__ X) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * kIn) return;
  const int l = t / kIn, f = t % kIn, s = f / kCh, c = f % kCh;
  const int o = nb9[9 * l + s];
  double v;
  if (c < 4)
    v = res[(size_t)c * nb + o] * nrm[0];
  else
    v = fprev[(size_t)(c - 4) * nb + o] * nrm[1];
  X[t] = (float)v / iscale[c];
}

// fb[c*nb + l] = fprev[c*nb + l] + Y[l][c] * oscale[c]   (c = 0..3)
__global__ void k_apply(int nb, const float* __restrict__ Y, const float* __restrict__ oscale,
                        const double* __restrict__ fprev, double* __restrict__ fb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * nb) return;
  const int c = t / nb, l = t % nb;
  const double d = (double)Y[(size_t)l * kOut + c] * (double)oscale[c];
  fb[t] = fprev ? fprev[t] + d : d;
}

template <bool TA, bool TB, int EPI, int T>
hipError_t gemm_tile(int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                     float* C, int ldc, const float* bias, const float* H, int ldh, hipStream_t s) {
  // float4 operand loads: the contiguous extent and the leading dimension
  // multiples of 4, the base 16-byte aligned
  auto al = [](const float* p) { return ((uintptr_t)p & 15) == 0; };
  const bool va = al(A) && lda % 4 == 0 && (TA ? M : K) % 4 == 0;
  const bool vb = al(B) && ldb % 4 == 0 && (TB ? K : N) % 4 == 0;
  const dim3 grid((M + T - 1) / T, (N + T - 1) / T);
  if (va && vb)
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, T, true, true>), grid, dim3(kBlock), 0, s, M, N, K, A,
                       lda, B, ldb, C, ldc, bias, H, ldh);
  else if (vb)
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, T, false, true>), grid, dim3(kBlock), 0, s, M, N, K, A,
                       lda, B, ldb, C, ldc, bias, H, ldh);
  else if (va)
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, T, true, false>), grid, dim3(kBlock), 0, s, M, N, K, A,
                       lda, B, ldb, C, ldc, bias, H, ldh);
  else
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, T, false, false>), grid, dim3(kBlock), 0, s, M, N, K, A,
                       lda, B, ldb, C, ldc, bias, H, ldh);
  return hipGetLastError();
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, const float* bias, const float* H, int ldh, hipStream_t s) {
  // 64 x 64 tiles when they give at least one workgroup per CU (256), else
  // 32 x 32 (the training batches: 256 x 256 products)
  const long long big = (long long)((M + 63) / 64) * ((N + 63) / 64);
  if (big >= 256) return gemm_tile<TA, TB, EPI, 64>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, s);
  return gemm_tile<TA, TB, EPI, 32>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, s);
}

}  // namespace

Layout layout() {
  Layout L{};
  const int dims[kLayers + 1] = {kIn, kHidden, kHidden, kHidden, kOut};
  int off = 0;
  for (int i = 0; i < kLayers; i++) {
    L.w[i] = off;
    off += dims[i + 1] * dims[i];
    L.b[i] = off;
    off += dims[i + 1];
    L.in[i] = dims[i];
    L.out[i] = dims[i + 1];
  }
  L.total = off;
  return L;
}

hipError_t forward(const Layout& L, const float* P, const float* X, int n, float* const act[kLayers],
                   hipStream_t s) {
  const float* in = X;
  for (int i = 0; i < kLayers; i++) {
    hipError_t e;
    if (i < kLayers - 1)
      e = gemm<false, true, kEpiBiasRelu>(n, L.out[i], L.in[i], in, L.in[i], P + L.w[i], L.in[i],
                                          act[i], L.out[i], P + L.b[i], nullptr, 0, s);
    else
      e = gemm<false, true, kEpiBias>(n, L.out[i], L.in[i], in, L.in[i], P + L.w[i], L.in[i],
                                      act[i], L.out[i], P + L.b[i], nullptr, 0, s);
    if (e != hipSuccess) return e;
    in = act[i];
  }
  return hipSuccess;
}

hipError_t loss_grad(const float* Y, const float* T, int n, float* dY, float* part, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(kLossBlocks), dim3(kBlock), 0, s, Y, T, n * kOut, dY, part);
  return hipGetLastError();
}

hipError_t backward(const Layout& L, const float* P, const float* X, int n,
                    float* const act[kLayers], float* dY, float* const dH[kLayers - 1], float* G,
                    hipStream_t s) {
  // layer i maps act[i-1] (X for i = 0) to act[i]; d holds dL/d(pre-activation)
  const float* d = dY;
  for (int i = kLayers - 1; i >= 0; i--) {
    const float* in = (i == 0) ? X : act[i - 1];
    hipError_t e = gemm<true, false, kEpiNone>(L.out[i], L.in[i], n, d, L.out[i], in, L.in[i],
                                               G + L.w[i], L.in[i], nullptr, nullptr, 0, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_colsum, dim3((L.out[i] + kColN - 1) / kColN), dim3(kBlock), 0, s, d, n,
                       L.out[i], G + L.b[i]);
    if (i > 0) {
      e = gemm<false, false, kEpiMask>(n, L.in[i], L.out[i], d, L.out[i], P + L.w[i], L.in[i],
                                       dH[i - 1], L.in[i], nullptr, act[i - 1], L.in[i], s);
      if (e != hipSuccess) return e;
      d = dH[i - 1];
    }
  }
  return hipGetLastError();
}

hipError_t adam(float* P, const float* G, float* M, float* V, int n, float lr_t, float b1, float b2,
                float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, G, M, V, n,
                     lr_t, b1, b2, eps);
  return hipGetLastError();
}

hipError_t features(int nb, const int* nb9, const double* res, const double* fprev,
                    const double* nrm, const float* iscale, float* X, hipStream_t s) {
  const size_t n = (size_t)nb * kIn;
  hipLaunchKernelGGL(k_features, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nb,
                     nb9, res, fprev, nrm, iscale, X);
  return hipGetLastError();
}

hipError_t apply(int nb, const float* Y, const float* oscale, const double* fprev, double* fb,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3((unsigned)((4 * (size_t)nb + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, s, nb, Y, oscale, fprev, fb);
  return hipGetLastError();
}

hipError_t gemm_test(int M, int N, int K, int ta, int tb, const float* A, int lda, const float* B,
                     int ldb, float* C, int ldc, hipStream_t s) {
  if (!ta && !tb) return gemm<false, false, kEpiNone>(M, N, K, A, lda, B, ldb, C, ldc, nullptr, nullptr, 0, s);
  if (!ta && tb) return gemm<false, true, kEpiNone>(M, N, K, A, lda, B, ldb, C, ldc, nullptr, nullptr, 0, s);
  if (ta && !tb) return gemm<true, false, kEpiNone>(M, N, K, A, lda, B, ldb, C, ldc, nullptr, nullptr, 0, s);
  return gemm<true, true, kEpiNone>(M, N, K, A, lda, B, ldb, C, ldc, nullptr, nullptr, 0, s);
}

}  // namespace mlp
}  // namespace c3d
