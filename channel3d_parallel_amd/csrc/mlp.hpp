// Learned initial guess for the IB Schur solve: network layout and launchers
// (mlp.hip).  Parameters live in one FP32 buffer: for each layer i the
// weight matrix W_i[out][in] (row-major) then the bias b_i[out].
#pragma once

#include <hip/hip_runtime.h>

namespace c3d {
namespace mlp {

constexpr int kCh = 6;           // per-marker channels
constexpr int kIn = 9 * kCh;     // 3x3 marker neighbourhood
constexpr int kHidden = 256;
constexpr int kOut = 4;          // dudn_x, dudn_y, dudn_z, p_jump
constexpr int kLayers = 4;
constexpr int kLossBlocks = 64;

constexpr int kEpiNone = 0, kEpiBias = 1, kEpiBiasRelu = 2, kEpiMask = 3;

struct Layout {
  int w[kLayers], b[kLayers], in[kLayers], out[kLayers];
  int total;
};
Layout layout();

// act[i] (n x out_i) = relu(act[i-1] W_i^T + b_i), the last layer linear.
hipError_t forward(const Layout& L, const float* P, const float* X, int n, float* const act[kLayers],
                   hipStream_t s);
// dY = dMSE/dY over n x 4, per-block loss partials (kLossBlocks).
hipError_t loss_grad(const float* Y, const float* T, int n, float* dY, float* part, hipStream_t s);
// G = dLoss/dP given dY; dH[i] (n x out_i) is scratch for the hidden layers.
hipError_t backward(const Layout& L, const float* P, const float* X, int n,
                    float* const act[kLayers], float* dY, float* const dH[kLayers - 1], float* G,
                    hipStream_t s);
hipError_t adam(float* P, const float* G, float* M, float* V, int n, float lr_t, float b1, float b2,
                float eps, hipStream_t s);
// IB feature rows X (nb x 54) from the warm start's Schur residual
// res = b - S fprev (4nb) and the previous force; nb9 holds 9 neighbour
// indices per marker.  The residual channels are multiplied by nrm[0] and
// the force channels by nrm[1] (device scalars: the solve's own 1/rms of
// each), then divided by iscale per channel.
hipError_t features(int nb, const int* nb9, const double* res, const double* fprev,
                    const double* nrm, const float* iscale, float* X, hipStream_t s);
// fb = fprev + Y * oscale per output channel (fprev == null: fb = Y * oscale).
hipError_t apply(int nb, const float* Y, const float* oscale, const double* fprev, double* fb,
                 hipStream_t s);
// Plain GEMM with the MFMA kernel (tests of the operand layout).
hipError_t gemm_test(int M, int N, int K, int ta, int tb, const float* A, int lda, const float* B,
                     int ldb, float* C, int ldc, hipStream_t s);

}  // namespace mlp
}  // namespace c3d
