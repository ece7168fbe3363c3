// C ABI of the channel3D projection library: context, grid set-up and the
// entry points declared in include/c3d.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/c3d.h"
#include "fft_lds.hpp"
#include "ib.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

using c3d::cpx;
using c3d::DevGrid;

// Learned-guess network of a context (allocated by c3d_mlp_init): FP32
// parameters, gradients and Adam moments, per-channel input/output scales,
// activation buffers sized for `cap` rows.
struct MLPState {
  c3d::mlp::Layout L{};
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  float *iscale = nullptr, *oscale = nullptr, *ones = nullptr;
  int t = 0;
  int cap = 0;
  float *X = nullptr, *T = nullptr, *dY = nullptr, *part = nullptr;
  float* act[c3d::mlp::kLayers] = {nullptr};
  float* dH[c3d::mlp::kLayers - 1] = {nullptr};
  std::vector<void*> bufs;  // activation buffers (freed on regrow)
  int* nb9 = nullptr;       // IB marker neighbourhoods
  double* nrm = nullptr;    // per-solve input normalisation (1/rms of residual, of force)
  bool ready = false;       // weights set by training or c3d_mlp_set_params
};

// Immersed-boundary state of a context (allocated by c3d_ib_setup).
struct IBState {
  c3d::ib::Markers m{};
  c3d::ib::Spread sp[3]{};
  // spread targets: zero except at support nodes and their ghost copies
  double *Ureg = nullptr, *Vreg = nullptr, *Wreg = nullptr;
  double* phi = nullptr;     // rhs_p-shaped Poisson solution / dt
  double* Hc_ext = nullptr;  // rhs_p-shaped exterior Heaviside
  double *Hu_int = nullptr, *Hv_int = nullptr, *Hw_int = nullptr;
  double *E1nHc = nullptr, *E1nH = nullptr;
  double *Uint = nullptr, *Vint = nullptr, *Wint = nullptr;  // U_interim ...
  double* aux = nullptr;  // 3nb
  double* vec[8] = {nullptr};  // x r rhat p v h s t  (4nb each)
  double *rhs_ib = nullptr, *fb = nullptr, *sc = nullptr, *part = nullptr, *mpart = nullptr;
  double *fin = nullptr, *fout = nullptr;  // 4nb staging for host-array calls
  double* fprev = nullptr;  // force before the last solve (learned-guess pairs)
  // converged force of each RK3 stage of the last step (guess mode 5,
  // allocated on first use): the start of the same stage of the next step
  double* fstage[3] = {nullptr};
  bool fstage_ok[3] = {false, false, false};
  // history-projection start (guess_mode 4), allocated on first use: ring of
  // the last kHist converged forces and their right-hand sides, and the
  // orthonormal rhs basis hQ with forces hX (S hX = hQ) built from it
  double* hF[c3d::ib::kHist] = {nullptr};
  double* hB[c3d::ib::kHist] = {nullptr};
  double* hQ[c3d::ib::kHist] = {nullptr};
  double* hX[c3d::ib::kHist] = {nullptr};
  double *hpart = nullptr, *hcoef = nullptr;
  int hcount = 0, hhead = 0;
  double hdt = 0.0;  // dt of the stored pairs (the Schur operator depends on dt)
  // One BiCGSTAB iteration (4 dots, 4 scalar steps, 3 vector updates, 2 Schur
  // products of ~16 kernels each, the 8-byte residual read-back) captured as a
  // HIP graph per (x, dt, rk_step == 1) and replayed once per iteration.
  struct IterGraph {
    const double* x;
    double dt;
    bool rk1;
    hipGraphExec_t exec;
  };
  std::vector<IterGraph> graphs;
  bool use_graphs = true;
  double* herr = nullptr;  // pinned host word: squared residual norm of the last iteration
  ~IBState() {
    for (auto& gr : graphs) (void)hipGraphExecDestroy(gr.exec);
    if (herr) (void)hipHostFree(herr);
  }
  // marker arrays (device) and the grid the stencil set-up reads
  double *xb = nullptr, *yb = nullptr, *zb = nullptr, *normals = nullptr, *sb = nullptr,
         *ub = nullptr;
  int* yref = nullptr;
  int* flags = nullptr;  // [0] geometry and [1] stencil range flags, [2..4] node counts
  c3d::ib::SetupGrid sg{};
  // moving wall body (c3d_ib_setup_wave): re-set-up at every Runge-Kutta stage
  bool moving = false;
  c3d::ib::WaveBody wave{};
  double ramp = 0.0;      // body_ramp_up_time
  double t_geom = 0.0;    // time of the current marker positions
  void* map_work = nullptr;  // scratch of the device map construction
  double last_dt = 1.0;     // dt of the last IB projection
  double tol = 1e-10;
  int maxit = 1000, guess = 0;
  int last_iters = 0;
  long long accum_iters = 0;
  double last_err = 0.0;
  bool ext_pg = false;
};

struct c3d_ctx {
  c3d_desc desc{};
  int nx = 0, ny = 0, nz = 0, nxg = 0, nyg = 0, nzg = 0, nxp = 0, nzp = 0, nym = 0, nxh = 0;
  // host grid (Fortran index i stored at [i-1])
  std::vector<double> x, y, z, xg, yg, zg, kxx, kzz, dyy_d, dyy_l, dyy_u;
  double dx = 0, dz = 0, dxmin = 0, dymin = 0, dzmin = 0, Lx = 0, Ly = 0, Lz = 0, Lxp = 0, Lzp = 0;
  int n_uniform = 1;  // uniform buffer points per wall (grid_type 2)
  double t = 0.0;     // simulation time (t of src/time_integration.f90)
  // device state
  hipStream_t stream = nullptr;
  DevGrid g{};
  std::vector<void*> allocs;
  double *U = nullptr, *V = nullptr, *W = nullptr, *P = nullptr, *Pint = nullptr, *rhs = nullptr;
  // staging of the host-array procedures (apply_boundary_conditions(U_,V_,W_),
  // divergence, gradient, regT, compute_rhs_u/v/w, the mass-flow integral):
  // field-shaped xU, xV, xW and the three right-hand-side interiors in xF,
  // allocated with the context, so those calls neither allocate nor touch
  // the resident state
  double *xU = nullptr, *xV = nullptr, *xW = nullptr, *xF = nullptr;
  cpx<double>* S = nullptr;
  double* CP = nullptr;
  unsigned long long* red = nullptr;
  unsigned* nanflag = nullptr;  // g.nonfinite: set by the velocity corrections
  double* psum = nullptr;  // 7*nyg plane sums of compute_statistics
  double* mon = nullptr;   // output_monitor partials and results
  size_t nu = 0, nv = 0, nw = 0, np = 0, nr = 0, nspec = 0;
  // RK3 state (allocated on first use): step-start copies, stage
  // right-hand sides F[stage][component], reduction scratch
  double *Uo = nullptr, *Vo = nullptr, *Wo = nullptr;
  // third set of velocity fields for the fused RK3 stages, which write each
  // stage's velocity into a set other than their input (the resident U, V,
  // W pointers then move among U, Uo and Ux)
  double *Ux = nullptr, *Vx = nullptr, *Wx = nullptr;
  bool rhs_exact = false;  // c3d_set_rhs_exact: bitwise right-hand sides in the time step
  double* F[3][3] = {{nullptr}};
  double* part = nullptr;  // mass-flow row partials, then 2 result slots
  double dPdx = 0.0, dPdy = 0.0, dPdz = 0.0;
  // interior Heaviside masks of the right-hand sides (null: H = 1) and of
  // the bulk-flow integral
  const double *Hu = nullptr, *Hv = nullptr, *Hw = nullptr, *Hm = nullptr;
  IBState* ib = nullptr;
  struct MLPState* mlp = nullptr;
  // refinement (nref sweeps per projection; mixed precision by default, FP64
  // contexts on request): FP64 divergence b, the first solve's potential
  // phi64 and the accumulated correction phic (kept apart: their FP64 sum
  // would round phi64's low bits away), the FP64 residual r64 of the first
  // sweep (when nref >= 2), gauge-row scratch; mixed: FP32 spectrum S32
  cpx<float>* S32 = nullptr;
  double *b64 = nullptr, *phi64 = nullptr, *phic = nullptr, *r64 = nullptr, *gauge = nullptr;
  int nref = 0;
  bool mixed_inv64 = true;  // FP64 inverse half of every mixed pass (C3D_MIXED_INV64=0: FP32)
  // modes whose tridiagonal systems are solved with FP64 arithmetic in the
  // mixed path: |kxx + kzz| < tri_thr (the (0,0) mode and the lowest |k|);
  // the others use FP32 arithmetic.  tri_low == null: FP64 for every mode.
  int* tri_low = nullptr;
  int ntri_low = 0;
  double tri_thr = 0.0;
  std::string err;
};

namespace {

int fail(c3d_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(ctx, C3D_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                  __FILE__, __LINE__);                                                        \
  } while (0)

template <typename T>
int dalloc(c3d_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 1);
  if (e != hipSuccess) {
    // clear the runtime's last-error slot, which the kernel launchers read
    (void)hipGetLastError();
    return fail(c, C3D_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T),
                hipGetErrorString(e));
  }
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return 0;
}

template <typename T>
int upload(c3d_ctx* c, const T** dst, const std::vector<T>& v) {
  T* p = nullptr;
  int rc = dalloc(c, &p, v.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *dst = p;
  return 0;
}

// Centres with mirror ghosts, src/initialization.f90:173-199.
void centres(const std::vector<double>& f, std::vector<double>& g) {
  const size_t n = f.size();
  g.assign(n + 1, 0.0);
  for (size_t i = 0; i + 1 < n; i++) g[i + 1] = 0.5 * (f[i] + f[i + 1]);
  g[0] = g[1] - 2.0 * (g[1] - f[0]);
  g[n] = g[n - 1] + 2.0 * (f[n - 1] - g[n - 1]);
}

double min_diff(const std::vector<double>& a) {
  double m = a[1] - a[0];
  for (size_t i = 1; i + 1 < a.size(); i++) m = std::fmin(m, a[i + 1] - a[i]);
  return m;
}

// Modified wavenumbers 2(cos(2 pi i/n) - 1)/d^2, src/initialization.f90:290-305
// (dx**2d0 evaluated as dx*dx).
std::vector<double> modified_wavenumbers(int n, double d) {
  const double pi = 4.0 * std::atan(1.0);  // src/global.f90:34
  std::vector<double> k(n, 0.0);
  const int half = (n + 1) / 2;
  for (int i = 0; i <= half && i < n; i++)
    k[i] = 2.0 * (std::cos(2.0 * pi * (double)i / (double)n) - 1.0) / (d * d);
  for (int i = half + 1; i < n; i++)
    k[i] = 2.0 * (std::cos(2.0 * pi * (double)(-n + i) / (double)n) - 1.0) / (d * d);
  return k;
}

template <typename real>
std::vector<cpx<real>> twiddles(int n) {
  std::vector<cpx<real>> t(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; m++) {
    const long double a = -2.0L * pi * (long double)m / (long double)n;
    t[m].x = (real)cosl(a);
    t[m].y = (real)sinl(a);
  }
  return t;
}

// src/input_output.f90:208-247: n_total points, the middle
// n_total - 2(n_uniform - 1) of them tanh-clustered on [-1, 1], scaled so
// that the clustered part and two half buffers span L; the n_uniform - 1
// points beyond each end continue at the smallest spacing.
void stretched_with_buffers(std::vector<double>& y, double L, int n_uniform, double alpha) {
  const int n = (int)y.size(), m = n - 2 * (n_uniform - 1);
  std::vector<double> s(m);
  for (int i = 0; i < m; i++) s[i] = (double)i;
  const double top = s[m - 1];
  for (int i = 0; i < m; i++) s[i] = 2.0 * s[i] / top - 1.0;
  if (alpha > 0.0)
    for (int i = 0; i < m; i++) s[i] = std::tanh(alpha * s[i]) / std::tanh(alpha);
  const double h_unit = s[1] - s[0];
  const double h = L * h_unit / (2.0 + (double)(n_uniform - 1) * h_unit);
  for (int i = 0; i < m; i++) y[n_uniform - 1 + i] = s[i] * h / h_unit;
  for (int b = 1; b < n_uniform; b++) {
    y[n - n_uniform + b] = y[n - n_uniform + b - 1] + h;
    y[n_uniform - 1 - b] = y[n_uniform - b] - h;
  }
}

// Grid generation, src/input_output.f90:140-206 and the derived quantities
// of src/initialization.f90:166-390.  Returns false when the type-2 buffers
// do not fit in ny points.
bool build_grid(c3d_ctx* c) {
  const c3d_desc& d = c->desc;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  c->x.resize(nx);
  c->z.resize(nz);
  c->y.resize(ny);
  for (int i = 0; i < nx; i++) c->x[i] = d.Lxp * (double)i / (double)(nx - 2);
  for (int i = 0; i < nz; i++) c->z[i] = d.Lzp * (double)i / (double)(nz - 2);
  if (d.grid_type == 0) {
    for (int i = 0; i < ny; i++) c->y[i] = d.Ly * (double)i / (double)(ny - 1);
  } else if (d.grid_type == 2) {
    // :180-203: grow the buffers one point at a time until they hold the
    // body amplitude plus the 2*suppy + 2 = 6 cells a delta stencil needs
    int nbuf = 1;
    for (;;) {
      stretched_with_buffers(c->y, d.Ly, nbuf, d.alpha);
      const double h = c->y[1] - c->y[0];
      if ((double)nbuf * h >= d.min_buffer_width + 6 * h) break;
      if (2 * ++nbuf >= ny) return false;
    }
    c->n_uniform = nbuf;
    for (int i = 0; i < ny; i++) c->y[i] += 1.0;
  } else {
    for (int i = 0; i < ny; i++) c->y[i] = d.Ly * (double)i / (double)(ny - 1) - 1.0;
    if (d.alpha > 0.0)
      for (int i = 0; i < ny; i++) c->y[i] = std::tanh(d.alpha * c->y[i]) / std::tanh(d.alpha);
    double mn = c->y[0];
    for (int i = 1; i < ny; i++) mn = std::fmin(mn, c->y[i]);
    for (int i = 0; i < ny; i++) c->y[i] = c->y[i] - mn;
    double mx = c->y[0];
    for (int i = 1; i < ny; i++) mx = std::fmax(mx, c->y[i]);
    for (int i = 0; i < ny; i++) c->y[i] = c->y[i] * d.Ly / mx;
  }
  centres(c->x, c->xg);
  centres(c->y, c->yg);
  centres(c->z, c->zg);
  c->dxmin = min_diff(c->xg);
  c->dymin = min_diff(c->yg);
  c->dzmin = min_diff(c->zg);
  c->Lx = c->x[nx - 1] - c->x[0];
  c->Ly = c->y[ny - 1] - c->y[0];
  c->Lz = c->z[nz - 1] - c->z[0];
  c->dx = c->dxmin;
  c->dz = c->dzmin;
  c->Lxp = c->Lx - c->dx;
  c->Lzp = c->Lz - c->dz;
  c->kxx = modified_wavenumbers(nx - 2, c->dx);
  c->kzz = modified_wavenumbers(nz - 2, c->dz);
  // Dyy, src/initialization.f90:353-383 (Fortran j = 2..nyg-1 -> jj = j-2)
  const int nyg = ny + 1, nym = ny - 1;
  c->dyy_d.assign(nym, 0.0);
  c->dyy_l.assign(nym > 1 ? nym - 1 : 1, 0.0);
  c->dyy_u.assign(nym > 1 ? nym - 1 : 1, 0.0);
  const std::vector<double>&y = c->y, &yg = c->yg;
  for (int j = 2; j <= nyg - 1; j++) {
    const double dyf = y[j - 1] - y[j - 2];
    const double a = 1.0 / dyf / (yg[j] - yg[j - 1]);
    const double b = 1.0 / dyf * (-1.0 / (yg[j] - yg[j - 1]) - 1.0 / (yg[j - 1] - yg[j - 2]));
    const double cc = 1.0 / dyf / (yg[j - 1] - yg[j - 2]);
    const int jj = j - 2;
    if (j == 2) {
      c->dyy_d[jj] = b + cc;
      c->dyy_u[jj] = a;
    } else if (j == nyg - 1) {
      c->dyy_d[jj] = a + b;
      c->dyy_l[jj - 1] = cc;
    } else {
      c->dyy_d[jj] = b;
      c->dyy_u[jj] = a;
      c->dyy_l[jj - 1] = cc;
    }
  }
  return true;
}

// Fortran-indexed spacing table: t[i] = a(i) - a(i-1) for i = 2..n (t[0], t[1] = 0).
std::vector<double> face_spacing(const std::vector<double>& a) {
  std::vector<double> t(a.size() + 2, 0.0);
  for (size_t i = 2; i <= a.size(); i++) t[i] = a[i - 1] - a[i - 2];
  return t;
}
// t[i] = a(i+1) - a(i) for i = 1..n-1.
std::vector<double> centre_spacing(const std::vector<double>& a) {
  std::vector<double> t(a.size() + 2, 0.0);
  for (size_t i = 1; i + 1 <= a.size(); i++) t[i] = a[i] - a[i - 1];
  return t;
}

// Interior sizes of the right-hand sides F_u (nx-2, nyg-2, nzg-2),
// F_v (nxg-2, ny-2, nzg-2), F_w (nxg-2, nyg-2, nz-2).
size_t interior_size(const c3d_ctx* c, int comp) {
  if (comp == 0) return (size_t)(c->nx - 2) * (c->nyg - 2) * (c->nzg - 2);
  if (comp == 1) return (size_t)(c->nxg - 2) * (c->ny - 2) * (c->nzg - 2);
  return (size_t)(c->nxg - 2) * (c->nyg - 2) * (c->nz - 2);
}
size_t interior_total(const c3d_ctx* c) {
  return interior_size(c, 0) + interior_size(c, 1) + interior_size(c, 2);
}

// Split of the mixed path's tridiagonal solve.  A mode's system Dyy + kxx +
// kzz has condition number about (max|Dyy| + |k2|)/|k2| with k2 = kxx + kzz;
// FP32 elimination is accurate to that many FP32 roundings, so the modes with
// |k2| < kTriLow * max|Dyy(j,j)| (condition above ~1e3, among them the
// singular (0,0) mode that the gauge row fixes) keep FP64 arithmetic.
// C3D_TRI32_LOW overrides the fraction; a negative value keeps FP64
// arithmetic for every mode.
constexpr double kTriLow = 1e-3;

int mixed_tri_modes(c3d_ctx* c) {
  static const double frac = [] {
    const char* e = getenv("C3D_TRI32_LOW");
    return e ? atof(e) : kTriLow;
  }();
  if (!(frac >= 0.0) || !c3d::tri_part_supported(c->g)) return 0;
  double dmax = 0.0;
  for (double v : c->dyy_d) dmax = std::max(dmax, std::fabs(v));
  // at least the smallest normal double, so that the (0,0) mode (k2 = 0) is
  // always in the FP64 set
  c->tri_thr = std::max(frac * dmax, 2.2250738585072014e-308);
  std::vector<int> low;
  for (int kx = 0; kx < c->nxh; kx++)
    for (int kz = 0; kz < c->nzp; kz++)
      if (std::fabs(c->kxx[kx] + c->kzz[kz]) < c->tri_thr) low.push_back(kx * c->nzp + kz);
  c->ntri_low = (int)low.size();
  const int* d = nullptr;
  int rc = upload(c, &d, low);
  c->tri_low = const_cast<int*>(d);
  return rc;
}

int sync(c3d_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// The Poisson pipeline between the forward and the inverse x transforms.
int spectral_core(c3d_ctx* c) {
  HIPCHK(c, c3d::launch_fft_z(c->g, -1, c->S, c->stream));
  HIPCHK(c, c3d::launch_tri(c->g, c->S, c->CP, c->stream));
  HIPCHK(c, c3d::launch_fft_z(c->g, +1, c->S, c->stream));
  return 0;
}

const char* kStageNames[] = {"fwd_x_div", "fft_z_fwd", "tri_y", "fft_z_inv", "inv_x",
                             "refine", "grad_correct"};
constexpr int kNumStages = 7;
constexpr int kMaxRef = 8;          // refinement sweeps per mixed projection
constexpr int kRedRefine = 4;       // first reduction slot of the sweep residual maxima
constexpr int kResSlots = 64;       // max slots per sweep (workgroup b uses slot b % 64)

// Moving-body re-set-up at time t (defined with the immersed-boundary set-up).
int ib_move_to(c3d_ctx* c, double t);

}  // namespace

extern "C" {

namespace {
// Device buffers of nref refinement sweeps, allocated on first need.
int refine_buffers(c3d_ctx* c, int nref) {
  int rc = 0;
  if (nref >= 1 && !c->b64 &&
      ((rc = dalloc(c, &c->b64, c->nr)) || (rc = dalloc(c, &c->phi64, c->nr)) ||
       (rc = dalloc(c, &c->phic, c->nr)) || (rc = dalloc(c, &c->gauge, (size_t)c->nzp + 1))))
    return rc;
  if (nref >= 2 && !c->r64 && (rc = dalloc(c, &c->r64, c->nr))) return rc;
  return 0;
}
}  // namespace

int c3d_create(const c3d_desc* desc, c3d_ctx** out) {
  if (!desc || !out) return C3D_EINVAL;
  *out = nullptr;
  const c3d_desc& d = *desc;
  // src/initialization.f90:46-47: even nx, nz; FFT lengths must be supported.
  if (d.nx < 6 || d.nz < 6 || d.ny < 4 || (d.nx % 2) || (d.nz % 2)) return C3D_EINVAL;
  if (!c3d::fft_length_supported(d.nx - 2) || !c3d::fft_length_supported(d.nz - 2))
    return C3D_EINVAL;
  if (d.grid_type < 0 || d.grid_type > 2) return C3D_EINVAL;
  if (d.grid_type == 2 && !(d.min_buffer_width >= 0.0)) return C3D_EINVAL;
  if (d.precision != C3D_PREC_FP64 && d.precision != C3D_PREC_MIXED) return C3D_EINVAL;
  if (d.precision == C3D_PREC_MIXED && d.ny - 1 != 64 && d.ny - 1 != 128 && d.ny - 1 != 192)
    return C3D_EINVAL;  // FP32-storage tridiagonal lengths
  if (!(d.Lxp > 0) || !(d.Lzp > 0) || !(d.Ly > 0)) return C3D_EINVAL;

  c3d_ctx* c = new c3d_ctx();
  c->desc = d;
  c->nx = d.nx; c->ny = d.ny; c->nz = d.nz;
  c->nxg = d.nx + 1; c->nyg = d.ny + 1; c->nzg = d.nz + 1;
  c->nxp = d.nx - 2; c->nzp = d.nz - 2; c->nym = d.ny - 1; c->nxh = c->nxp / 2 + 1;
  if (!build_grid(c)) {  // type-2 buffers exceed the grid (src/input_output.f90:199)
    delete c;
    return C3D_EINVAL;
  }

  auto cleanup = [&](int rc) {
    c3d_destroy(c);
    return rc;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(C3D_EHIP);

  DevGrid& g = c->g;
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
  g.nxg = c->nxg; g.nyg = c->nyg; g.nzg = c->nzg;
  g.nxp = c->nxp; g.nzp = c->nzp; g.nym = c->nym; g.nxh = c->nxh;
  int rc = 0;
  auto recip = [](std::vector<double> v) {
    for (double& e : v) e = (e != 0.0) ? 1.0 / e : 0.0;
    return v;
  };
  // nu/h with the quotient rounded once, as nu*1.0/h evaluates
  auto nu_over = [&](std::vector<double> v) {
    for (double& e : v) e = (e != 0.0) ? (d.nu * 1.0) / e : 0.0;
    return v;
  };
  if ((rc = upload(c, &g.rdxf, recip(face_spacing(c->x)))) ||
      (rc = upload(c, &g.rdyf, recip(face_spacing(c->y)))) ||
      (rc = upload(c, &g.rdzf, recip(face_spacing(c->z)))) ||
      (rc = upload(c, &g.rdxg, recip(centre_spacing(c->xg)))) ||
      (rc = upload(c, &g.rdyg, recip(centre_spacing(c->yg)))) ||
      (rc = upload(c, &g.rdzg, recip(centre_spacing(c->zg)))) ||
      (rc = upload(c, &g.nudxf, nu_over(face_spacing(c->x)))) ||
      (rc = upload(c, &g.nudyf, nu_over(face_spacing(c->y)))) ||
      (rc = upload(c, &g.nudzf, nu_over(face_spacing(c->z)))) ||
      (rc = upload(c, &g.nudxg, nu_over(centre_spacing(c->xg)))) ||
      (rc = upload(c, &g.nudzg, nu_over(centre_spacing(c->zg)))))
    return cleanup(rc);
  if ((rc = upload(c, &g.dxf, face_spacing(c->x))) || (rc = upload(c, &g.dyf, face_spacing(c->y))) ||
      (rc = upload(c, &g.dzf, face_spacing(c->z))) ||
      (rc = upload(c, &g.dxg, centre_spacing(c->xg))) ||
      (rc = upload(c, &g.dyg, centre_spacing(c->yg))) ||
      (rc = upload(c, &g.dzg, centre_spacing(c->zg))) || (rc = upload(c, &g.kxx, c->kxx)) ||
      (rc = upload(c, &g.kzz, c->kzz)) || (rc = upload(c, &g.dyy_d, c->dyy_d)) ||
      (rc = upload(c, &g.dyy_l, c->dyy_l)) || (rc = upload(c, &g.dyy_u, c->dyy_u)) ||
      (rc = upload(c, &g.twx64, twiddles<double>(c->nxp))) ||
      (rc = upload(c, &g.twz64, twiddles<double>(c->nzp))) ||
      (rc = upload(c, &g.twx32, twiddles<float>(c->nxp))) ||
      (rc = upload(c, &g.twz32, twiddles<float>(c->nzp))))
    return cleanup(rc);

  c->nu = (size_t)c->nx * c->nyg * c->nzg;
  c->nv = (size_t)c->nxg * c->ny * c->nzg;
  c->nw = (size_t)c->nxg * c->nyg * c->nz;
  c->np = (size_t)c->nxg * c->nyg * c->nzg;
  c->nr = (size_t)(c->nxg - 2) * (c->nyg - 2) * (c->nzg - 1);
  c->nspec = (size_t)c->nym * c->nxh * c->nzp;
  if ((rc = dalloc(c, &c->U, c->nu)) || (rc = dalloc(c, &c->V, c->nv)) ||
      (rc = dalloc(c, &c->W, c->nw)) || (rc = dalloc(c, &c->P, c->np)) ||
      (rc = dalloc(c, &c->Pint, c->nr)) || (rc = dalloc(c, &c->rhs, c->nr)) ||
      (rc = dalloc(c, &c->xU, c->nu)) || (rc = dalloc(c, &c->xV, c->nv)) ||
      (rc = dalloc(c, &c->xW, c->nw)) || (rc = dalloc(c, &c->xF, interior_total(c))) ||
      (rc = dalloc(c, &c->part, 2 * (size_t)c3d::massflow_rows(c->g) + 2)) ||
      (rc = dalloc(c, &c->nanflag, 1)) ||
      (rc = dalloc(c, &c->S, c->nspec)) ||
      (!c3d::tri_part_supported(c->g) && (rc = dalloc(c, &c->CP, c->nspec))) ||
      (rc = dalloc(c, &c->red, kRedRefine + (size_t)kMaxRef * kResSlots)))
    return cleanup(rc);
  if (d.precision == C3D_PREC_MIXED && (rc = dalloc(c, &c->S32, c->nspec))) return cleanup(rc);
  c->g.nonfinite = c->nanflag;
  {
    const char* e = getenv("C3D_RHS_EXACT");
    c->rhs_exact = e && atoi(e) != 0;
  }
  if (d.precision == C3D_PREC_MIXED) {
    if ((rc = mixed_tri_modes(c))) return cleanup(rc);
    const char* e = getenv("C3D_MIXED_INV64");
    c->mixed_inv64 = !(e && atoi(e) == 0);
    // sweeps that bring max|div| below 1e-10 at 384x193x384 (one sweep
    // leaves ~1e-14 of the input's max|div|; two reach the FP64 floor)
    c->nref = c->mixed_inv64 ? 1 : 2;
    if ((rc = refine_buffers(c, c->nref))) return cleanup(rc);
  }
  if (hipMemsetAsync(c->nanflag, 0, sizeof(unsigned), c->stream) != hipSuccess ||
      hipMemsetAsync(c->U, 0, c->nu * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(c->V, 0, c->nv * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(c->W, 0, c->nw * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(c->P, 0, c->np * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(c->Pint, 0, c->nr * 8, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return cleanup(C3D_EHIP);
  *out = c;
  return C3D_OK;
}

int c3d_destroy(c3d_ctx* c) {
  if (!c) return C3D_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->mlp)
    for (void* p : c->mlp->bufs) (void)hipFree(p);
  delete c->mlp;
  delete c->ib;
  delete c;
  return C3D_OK;
}

const char* c3d_last_error(const c3d_ctx* c) { return c ? c->err.c_str() : "null context"; }

int c3d_grid_get(const c3d_ctx* c, int what, double* out, int n) {
  if (!c || !out) return C3D_EINVAL;
  const std::vector<double>* v = nullptr;
  std::vector<double> s;
  switch (what) {
    case 0: v = &c->x; break;
    case 1: v = &c->y; break;
    case 2: v = &c->z; break;
    case 3: v = &c->xg; break;
    case 4: v = &c->yg; break;
    case 5: v = &c->zg; break;
    case 6: v = &c->kxx; break;
    case 7: v = &c->kzz; break;
    case 8: v = &c->dyy_d; break;
    case 9: v = &c->dyy_l; break;
    case 10: v = &c->dyy_u; break;
    case 11:
      s = {c->dx, c->dz, c->dxmin, c->dymin, c->dzmin, c->Lx, c->Ly, c->Lz, c->Lxp, c->Lzp};
      v = &s;
      break;
    case 12:
      s = {(double)c->n_uniform};
      v = &s;
      break;
    default: return C3D_EINVAL;
  }
  if ((int)v->size() > n) return C3D_EINVAL;
  std::memcpy(out, v->data(), v->size() * sizeof(double));
  return (int)v->size();
}

int c3d_set_fields(c3d_ctx* c, const double* U, const double* V, const double* W, const double* P) {
  if (!c) return C3D_EINVAL;
  if (U) HIPCHK(c, hipMemcpyAsync(c->U, U, c->nu * 8, hipMemcpyHostToDevice, c->stream));
  if (V) HIPCHK(c, hipMemcpyAsync(c->V, V, c->nv * 8, hipMemcpyHostToDevice, c->stream));
  if (W) HIPCHK(c, hipMemcpyAsync(c->W, W, c->nw * 8, hipMemcpyHostToDevice, c->stream));
  if (P) HIPCHK(c, hipMemcpyAsync(c->P, P, c->np * 8, hipMemcpyHostToDevice, c->stream));
  return sync(c);
}

int c3d_get_fields(c3d_ctx* c, double* U, double* V, double* W, double* P, double* Pint) {
  if (!c) return C3D_EINVAL;
  if (U) HIPCHK(c, hipMemcpyAsync(U, c->U, c->nu * 8, hipMemcpyDeviceToHost, c->stream));
  if (V) HIPCHK(c, hipMemcpyAsync(V, c->V, c->nv * 8, hipMemcpyDeviceToHost, c->stream));
  if (W) HIPCHK(c, hipMemcpyAsync(W, c->W, c->nw * 8, hipMemcpyDeviceToHost, c->stream));
  if (P) HIPCHK(c, hipMemcpyAsync(P, c->P, c->np * 8, hipMemcpyDeviceToHost, c->stream));
  if (Pint) HIPCHK(c, hipMemcpyAsync(Pint, c->Pint, c->nr * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_apply_boundary_conditions(c3d_ctx* c) {
  if (!c) return C3D_EINVAL;
  HIPCHK(c, c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream));
  return sync(c);
}

int c3d_divergence(c3d_ctx* c, double* div) {
  if (!c || !div) return C3D_EINVAL;
  HIPCHK(c, c3d::launch_divergence(c->g, c->U, c->V, c->W, c->rhs, c->stream));
  HIPCHK(c, hipMemcpyAsync(div, c->rhs, c->nr * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_gradient(c3d_ctx* c, const double* p, double* Gx, double* Gy, double* Gz) {
  if (!c || !p || !Gx || !Gy || !Gz) return C3D_EINVAL;
  HIPCHK(c, hipMemcpyAsync(c->rhs, p, c->nr * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->xU, 0, c->nu * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->xV, 0, c->nv * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->xW, 0, c->nw * 8, c->stream));
  HIPCHK(c, c3d::launch_gradient(c->g, c->rhs, c->xU, c->xV, c->xW, c->stream));
  HIPCHK(c, hipMemcpyAsync(Gx, c->xU, c->nu * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Gy, c->xV, c->nv * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Gz, c->xW, c->nw * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

// Stages host arrays U_, V_, W_ (any may be NULL) in the context's scratch.
static int stage_uvw(c3d_ctx* c, const double* U, const double* V, const double* W) {
  if (U) HIPCHK(c, hipMemcpyAsync(c->xU, U, c->nu * 8, hipMemcpyHostToDevice, c->stream));
  if (V) HIPCHK(c, hipMemcpyAsync(c->xV, V, c->nv * 8, hipMemcpyHostToDevice, c->stream));
  if (W) HIPCHK(c, hipMemcpyAsync(c->xW, W, c->nw * 8, hipMemcpyHostToDevice, c->stream));
  return 0;
}

int c3d_apply_boundary_conditions_arrays(c3d_ctx* c, double* U, double* V, double* W) {
  if (!c || !U || !V || !W) return C3D_EINVAL;
  int rc = stage_uvw(c, U, V, W);
  if (rc) return rc;
  HIPCHK(c, c3d::launch_apply_bc(c->g, c->xU, c->xV, c->xW, c->stream));
  HIPCHK(c, hipMemcpyAsync(U, c->xU, c->nu * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(V, c->xV, c->nv * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(W, c->xW, c->nw * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_divergence_arrays(c3d_ctx* c, const double* U, const double* V, const double* W,
                          double* div) {
  if (!c || !U || !V || !W || !div) return C3D_EINVAL;
  int rc = stage_uvw(c, U, V, W);
  if (rc) return rc;
  HIPCHK(c, c3d::launch_divergence(c->g, c->xU, c->xV, c->xW, c->rhs, c->stream));
  HIPCHK(c, hipMemcpyAsync(div, c->rhs, c->nr * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_solve_poisson_equation(c3d_ctx* c, double* rhs, int rk_step, double dt) {
  if (!c || !rhs) return C3D_EINVAL;
  if (rk_step == 1 && !(dt != 0.0)) return fail(c, C3D_EINVAL, "dt must be nonzero at rk_step 1");
  HIPCHK(c, hipMemcpyAsync(c->rhs, rhs, c->nr * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, c3d::launch_fwd_x(c->g, false, nullptr, nullptr, nullptr, c->rhs, c->S, c->stream));
  int rc = spectral_core(c);
  if (rc) return rc;
  HIPCHK(c, c3d::launch_inv_x(c->g, 0, c->S, c->rhs, c->P, dt, rk_step == 1, c->stream));
  if (rk_step == 1) HIPCHK(c, c3d::launch_p_ghosts(c->g, c->P, c->stream));
  HIPCHK(c, hipMemcpyAsync(rhs, c->rhs, c->nr * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

// compute_non_IB_projection (src/projection.f90:22-46) on the resident
// velocity.  FP64 without refinement: one spectral solve, the reference's
// operations.  With refinement (mixed precision, or FP64 on request): a first
// solve for phi1, then nref sweeps that each transform the FP64 residual
// r = b - M (phi1 + phic) (M = D G plus the spectral solver's gauge row) and
// add the solution to phic; the velocity is corrected by G(phi1) + G(phic).
// Mixed: the forward half of every pass (x and z transforms) on FP32
// spectra, the inverse half (tridiagonal solution, inverse z and x
// transforms) in FP64 (or FP32 with C3D_MIXED_INV64=0).
static int projection_launch(c3d_ctx* c, double dt, int rk_step, hipEvent_t* ev) {
  const bool rk1 = rk_step == 1;
  const bool mixed = c->desc.precision == C3D_PREC_MIXED;
  const int nref = c->nref;
  const DevGrid& g = c->g;
  hipStream_t st = c->stream;
  if (ev) HIPCHK(c, hipEventRecord(ev[0], st));
  if (!mixed && nref == 0) {
    HIPCHK(c, c3d::launch_fwd_x(g, true, c->U, c->V, c->W, nullptr, c->S, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[1], st));
    HIPCHK(c, c3d::launch_fft_z(g, -1, c->S, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[2], st));
    HIPCHK(c, c3d::launch_tri(g, c->S, c->CP, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[3], st));
    HIPCHK(c, c3d::launch_fft_z(g, +1, c->S, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[4], st));
    HIPCHK(c, c3d::launch_inv_x(g, 1, c->S, c->Pint, c->P, dt, rk1, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[5], st));
    if (rk1) HIPCHK(c, c3d::launch_p_ghosts(g, c->P, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[6], st));
    HIPCHK(c, c3d::launch_grad_correct(g, c->U, c->V, c->W, c->Pint, dt, st));
    if (ev) HIPCHK(c, hipEventRecord(ev[7], st));
    return 0;
  }
  // the FP32 inverse half of the mixed variant
  const bool inv32 = mixed && !c->mixed_inv64;
  auto fwd_z = [&]() -> hipError_t {
    return mixed ? c3d::launch_fft_z32(g, -1, c->S32, st) : c3d::launch_fft_z(g, -1, c->S, st);
  };
  // inverse half of a pass: inv_x mode (0 phi, 1 phi/dt, 3 in1 + phi, 5 the
  // final potential, see launch_inv_x64_ref) into out (and out2)
  auto back = [&](int mode, const double* in1, const double* in2, double* out, double* out2,
                  bool rk1w, bool events) -> int {
    if (!mixed) HIPCHK(c, c3d::launch_tri(g, c->S, c->CP, st));
    else if (inv32) HIPCHK(c, c3d::launch_tri32(g, c->S32, c->tri_low, c->ntri_low, c->tri_thr, st));
    else HIPCHK(c, c3d::launch_tri32to64(g, c->S32, c->S, st));
    if (events && ev) HIPCHK(c, hipEventRecord(ev[3], st));
    HIPCHK(c, inv32 ? c3d::launch_fft_z32(g, +1, c->S32, st) : c3d::launch_fft_z(g, +1, c->S, st));
    if (events && ev) HIPCHK(c, hipEventRecord(ev[4], st));
    if (inv32)
      HIPCHK(c, c3d::launch_inv_x32(g, mode, c->S32, in1, in2, out, out2, c->P, dt, rk1w, st));
    else
      HIPCHK(c, c3d::launch_inv_x64_ref(g, mode, c->S, in1, in2, out, out2, c->P, dt, rk1w, st));
    return 0;
  };
  int rc = 0;
  if (nref > 0)
    HIPCHK(c, hipMemsetAsync(c->red + kRedRefine, 0,
                             (size_t)kMaxRef * kResSlots * sizeof(unsigned long long), st));
  // first solve; the FP64 divergence b is kept only for the refinement
  if (mixed)
    HIPCHK(c, c3d::launch_fwd_x32(g, c->U, c->V, c->W, nref > 0 ? c->b64 : nullptr, c->S32, st));
  else
    HIPCHK(c, c3d::launch_fwd_x_b64(g, c->U, c->V, c->W, c->b64, c->S, st));
  if (ev) HIPCHK(c, hipEventRecord(ev[1], st));
  HIPCHK(c, fwd_z());
  if (ev) HIPCHK(c, hipEventRecord(ev[2], st));
  if (nref == 0) rc = back(1, nullptr, nullptr, c->Pint, nullptr, rk1, true);
  else rc = back(0, nullptr, nullptr, c->phi64, nullptr, false, true);
  if (rc) return rc;
  if (ev) HIPCHK(c, hipEventRecord(ev[5], st));
  for (int s = 1; s <= nref; s++) {
    // sweep 1: r = b - M phi1 (kept in r64 for later sweeps); sweep s > 1:
    // r = r64 - M phic, M linear
    unsigned long long* rslots = c->red + kRedRefine + (size_t)(s - 1) * kResSlots;
    const double* bs = s == 1 ? c->b64 : c->r64;
    const double* ps = s == 1 ? c->phi64 : c->phic;
    double* rout = (s == 1 && nref >= 2) ? c->r64 : nullptr;
    if (mixed)
      HIPCHK(c, c3d::launch_fwd_x_resid32(g, bs, ps, c->gauge, c->S32, rslots, rout, st));
    else
      HIPCHK(c, c3d::launch_fwd_x_resid64(g, bs, ps, c->gauge, c->S, rslots, rout, st));
    HIPCHK(c, fwd_z());
    if (s < nref)
      rc = back(s == 1 ? 0 : 3, c->phic, nullptr, c->phic, nullptr, false, false);
    else
      rc = back(5, c->phi64, s == 1 ? nullptr : c->phic, nullptr, c->phic, rk1, false);
    if (rc) return rc;
  }
  if (rk1) HIPCHK(c, c3d::launch_p_ghosts(g, c->P, st));
  if (ev) HIPCHK(c, hipEventRecord(ev[6], st));
  if (nref == 0)
    HIPCHK(c, c3d::launch_grad_correct(g, c->U, c->V, c->W, c->Pint, dt, st));
  else
    HIPCHK(c, c3d::launch_grad_correct_split(g, c->U, c->V, c->W, c->phi64, c->phic, c->Pint, dt,
                                             st));
  if (ev) HIPCHK(c, hipEventRecord(ev[7], st));
  return 0;
}

// U_interim = U ... (src/projection.f90:26-28), kept only when a body needs it.
static int save_interim(c3d_ctx* c) {
  if (!c->ib) return 0;
  HIPCHK(c, hipMemcpyAsync(c->ib->Uint, c->U, c->nu * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ib->Vint, c->V, c->nv * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ib->Wint, c->W, c->nw * 8, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

static int ib_projection_launch(c3d_ctx* c, double dt, int rk_step);
static int mlp_guess(c3d_ctx* c, double dt, int rk_step);

int c3d_compute_non_ib_projection(c3d_ctx* c, double dt, int rk_step) {
  if (!c) return C3D_EINVAL;
  if (!(dt != 0.0)) return fail(c, C3D_EINVAL, "dt must be nonzero");
  int rc = save_interim(c);
  if (!rc) rc = projection_launch(c, dt, rk_step, nullptr);
  if (rc) return rc;
  return sync(c);
}

int c3d_check_divergence(c3d_ctx* c, double* max_div) {
  if (!c || !max_div) return C3D_EINVAL;
  HIPCHK(c, hipMemsetAsync(c->red, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, c3d::launch_max_div(c->g, c->U, c->V, c->W, c->red, c->stream));
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->red, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  std::memcpy(max_div, &bits, sizeof(double));
  return C3D_OK;
}

int c3d_copy_bandwidth(c3d_ctx* c, size_t nbytes, int iters, double* ms) {
  if (!c || !ms || iters <= 0 || nbytes < 16 || (nbytes % 16)) return C3D_EINVAL;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&a, nbytes) != hipSuccess || hipMalloc(&b, nbytes) != hipSuccess) {
    (void)hipGetLastError();  // the launchers report hipGetLastError()
    rc = fail(c, C3D_ENOMEM, "copy calibration buffers");
  } else if (hipMemsetAsync(a, 0, nbytes, c->stream) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
             c3d::launch_copy(a, b, nbytes, c->stream) != hipSuccess ||
             hipEventRecord(e0, c->stream) != hipSuccess) {
    rc = fail(c, C3D_EHIP, "copy calibration set-up failed");
  } else {
    for (int it = 0; it < iters && !rc; it++)
      if (c3d::launch_copy(a, b, nbytes, c->stream) != hipSuccess) rc = fail(c, C3D_EHIP, "copy");
    float t = 0.f;
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = fail(c, C3D_EHIP, "copy timing failed");
    if (!rc) *ms = t / iters;
  }
  (void)hipStreamSynchronize(c->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ----------------------------------------------------------------------------
// RK3 time integration (src/time_integration.f90, src/equations.f90,
// src/mass_flow.f90)
// ----------------------------------------------------------------------------
static int ensure_rk(c3d_ctx* c) {
  if (c->Uo) return 0;
  const size_t fu = (size_t)(c->nx - 2) * (c->nyg - 2) * (c->nzg - 2);
  const size_t fv = (size_t)(c->nxg - 2) * (c->ny - 2) * (c->nzg - 2);
  const size_t fw = (size_t)(c->nxg - 2) * (c->nyg - 2) * (c->nz - 2);
  int rc = 0;
  if ((rc = dalloc(c, &c->Uo, c->nu)) || (rc = dalloc(c, &c->Vo, c->nv)) ||
      (rc = dalloc(c, &c->Wo, c->nw)) || (rc = dalloc(c, &c->Ux, c->nu)) ||
      (rc = dalloc(c, &c->Vx, c->nv)) || (rc = dalloc(c, &c->Wx, c->nw))) {
    c->Uo = nullptr;  // allocate again next time
    return rc;
  }
  for (int s = 0; s < 3; s++)
    if ((rc = dalloc(c, &c->F[s][0], fu)) || (rc = dalloc(c, &c->F[s][1], fv)) ||
        (rc = dalloc(c, &c->F[s][2], fw)))
      return rc;
  for (int s = 0; s < 3; s++) {
    HIPCHK(c, hipMemsetAsync(c->F[s][0], 0, fu * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->F[s][1], 0, fv * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->F[s][2], 0, fw * 8, c->stream));
  }
  return 0;
}

static int save_state(c3d_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->Uo, c->U, c->nu * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->Vo, c->V, c->nv * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->Wo, c->W, c->nw * 8, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

static int rhs_launch(c3d_ctx* c, int slot) {
  double** F = c->F[slot - 1];
  HIPCHK(c, c3d::launch_rhs(c->g, c->U, c->V, c->W, c->Hu, c->Hv, c->Hw, c->desc.nu, c->dPdx,
                            c->dPdy, c->dPdz, F[0], F[1], F[2], c->stream));
  return 0;
}

static int rk_update_launch(c3d_ctx* c, int rk_step, double dt) {
  double* X[3] = {c->U, c->V, c->W};
  const double* Xo[3] = {c->Uo, c->Vo, c->Wo};
  for (int q = 0; q < 3; q++)
    HIPCHK(c, c3d::launch_rk_update(c->g, q, rk_step, dt, c->F[0][q], c->F[1][q], c->F[2][q],
                                    Xo[q], X[q], c->stream));
  return 0;
}

// One fused stage (right-hand side and update in one kernel, to round-off of
// the reference's operations).  Stage 1 reads U and writes into the Uo set,
// which then becomes U while the old U (the step's start) becomes Uo; stages
// 2 and 3 read U, Uo and F1 and write into the Ux set, which then swaps with U.
static int rk_stage_fast(c3d_ctx* c, int st, double dt) {
  c3d::RkStage a{};
  a.U = c->U;
  a.V = c->V;
  a.W = c->W;
  a.Hu = c->Hu;
  a.Hv = c->Hv;
  a.Hw = c->Hw;
  a.nu = c->desc.nu;
  a.dPdx = c->dPdx;
  a.dPdy = c->dPdy;
  a.dPdz = c->dPdz;
  a.dt = dt;
  a.Uo = c->Uo;
  a.Vo = c->Vo;
  a.Wo = c->Wo;
  a.F1u = c->F[0][0];
  a.F1v = c->F[0][1];
  a.F1w = c->F[0][2];
  if (st < 3) {
    a.Fu = c->F[st - 1][0];
    a.Fv = c->F[st - 1][1];
    a.Fw = c->F[st - 1][2];
  }
  double** dst[3] = {st == 1 ? &c->Uo : &c->Ux, st == 1 ? &c->Vo : &c->Vx, st == 1 ? &c->Wo : &c->Wx};
  a.Un = *dst[0];
  a.Vn = *dst[1];
  a.Wn = *dst[2];
  HIPCHK(c, c3d::launch_rhs_stage(c->g, st, a, c->stream));
  std::swap(c->U, *dst[0]);
  std::swap(c->V, *dst[1]);
  std::swap(c->W, *dst[2]);
  return 0;
}

static int mass_flow(c3d_ctx* c, const double* U, double* q) {
  double* out = c->part + 2 * (size_t)c3d::massflow_rows(c->g);
  HIPCHK(c, c3d::launch_massflow(c->g, U, c->Hm, c->part, out, c->stream));
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  *q = h[0] / h[1];
  return 0;
}

int c3d_set_pressure_gradient(c3d_ctx* c, double dPdx, double dPdy, double dPdz) {
  if (!c) return C3D_EINVAL;
  c->dPdx = dPdx;
  c->dPdy = dPdy;
  c->dPdz = dPdz;
  return C3D_OK;
}

int c3d_get_pressure_gradient(const c3d_ctx* c, double* dPdx, double* dPdy, double* dPdz) {
  if (!c) return C3D_EINVAL;
  if (dPdx) *dPdx = c->dPdx;
  if (dPdy) *dPdy = c->dPdy;
  if (dPdz) *dPdz = c->dPdz;
  return C3D_OK;
}

int c3d_save_state(c3d_ctx* c) {
  if (!c) return C3D_EINVAL;
  int rc = ensure_rk(c);
  if (!rc) rc = save_state(c);
  if (!rc) rc = sync(c);
  return rc;
}

int c3d_compute_rhs(c3d_ctx* c, int slot) {
  if (!c || slot < 1 || slot > 3) return C3D_EINVAL;
  int rc = ensure_rk(c);
  if (!rc) rc = rhs_launch(c, slot);
  if (!rc) rc = sync(c);
  return rc;
}

int c3d_get_rhs(c3d_ctx* c, int slot, double* Fu, double* Fv, double* Fw) {
  if (!c || slot < 1 || slot > 3) return C3D_EINVAL;
  int rc = ensure_rk(c);
  if (rc) return rc;
  const size_t fu = (size_t)(c->nx - 2) * (c->nyg - 2) * (c->nzg - 2);
  const size_t fv = (size_t)(c->nxg - 2) * (c->ny - 2) * (c->nzg - 2);
  const size_t fw = (size_t)(c->nxg - 2) * (c->nyg - 2) * (c->nz - 2);
  double** F = c->F[slot - 1];
  if (Fu) HIPCHK(c, hipMemcpyAsync(Fu, F[0], fu * 8, hipMemcpyDeviceToHost, c->stream));
  if (Fv) HIPCHK(c, hipMemcpyAsync(Fv, F[1], fv * 8, hipMemcpyDeviceToHost, c->stream));
  if (Fw) HIPCHK(c, hipMemcpyAsync(Fw, F[2], fw * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_rk3_update(c3d_ctx* c, int rk_step, double dt) {
  if (!c || rk_step < 1 || rk_step > 3) return C3D_EINVAL;
  int rc = ensure_rk(c);
  if (!rc) rc = rk_update_launch(c, rk_step, dt);
  if (!rc) rc = sync(c);
  return rc;
}

int c3d_compute_dt(c3d_ctx* c, double cfl, double* dt_out) {
  if (!c || !dt_out) return C3D_EINVAL;
  unsigned long long h[3] = {0, 0, 0};
  unsigned long long* red3 = c->red + 1;  // slots 1..3 of the reduction buffer
  HIPCHK(c, hipMemsetAsync(red3, 0, sizeof(h), c->stream));
  HIPCHK(c, c3d::launch_dt_max(c->g, c->U, c->V, c->W, red3, c->stream));
  HIPCHK(c, hipMemcpyAsync(h, red3, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  double lmax[3];
  std::memcpy(lmax, h, sizeof(lmax));
  // src/time_integration.f90:241-265
  const double dtc = std::fmin(std::fmin(cfl * lmax[0], cfl * lmax[1]), cfl * lmax[2]);
  const double nu = c->desc.nu;
  const double dtv = std::fmin(std::fmin(cfl * (c->dxmin * c->dxmin) / nu,
                                         cfl * (c->dymin * c->dymin) / nu),
                               cfl * (c->dzmin * c->dzmin) / nu);
  double dt = std::fmin(dtc, dtv);
  dt = std::fmin(dt, 0.1);
  if (cfl < 0) dt = -cfl;
  *dt_out = dt;
  return C3D_OK;
}

int c3d_mean_mass_flow_u(c3d_ctx* c, double* q) {
  if (!c || !q) return C3D_EINVAL;
  return mass_flow(c, c->U, q);
}

int c3d_mean_mass_flow_u_array(c3d_ctx* c, const double* U, double* q) {
  if (!c || !U || !q) return C3D_EINVAL;
  int rc = stage_uvw(c, U, nullptr, nullptr);
  if (!rc) rc = mass_flow(c, c->xU, q);
  return rc;
}

int c3d_compute_rhs_arrays(c3d_ctx* c, const double* U, const double* V, const double* W,
                           double* Fu, double* Fv, double* Fw) {
  if (!c || !U || !V || !W) return C3D_EINVAL;
  int rc = stage_uvw(c, U, V, W);
  if (rc) return rc;
  double* F[3] = {c->xF, c->xF + interior_size(c, 0), c->xF + interior_size(c, 0) + interior_size(c, 1)};
  HIPCHK(c, c3d::launch_rhs(c->g, c->xU, c->xV, c->xW, c->Hu, c->Hv, c->Hw, c->desc.nu, c->dPdx,
                            c->dPdy, c->dPdz, F[0], F[1], F[2], c->stream));
  double* out[3] = {Fu, Fv, Fw};
  for (int q = 0; q < 3; q++)
    if (out[q])
      HIPCHK(c, hipMemcpyAsync(out[q], F[q], interior_size(c, q) * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

// Plane means of compute_statistics (src/statistics.f90:41-111, one domain):
// means[q*nyg + j-1], q = 0 Umean, 1 Vmean, 2 Wmean, 3 U2mean, 4 V2mean,
// 5 W2mean, 6 UVmean; U, V, U2, V2, UV over (nxg-3)*(nzg-3) cells, W, W2 over
// (nxg-3)*(nz-2) (:103-111).
static int plane_means(c3d_ctx* c, std::vector<double>& m) {
  const int n = c->nyg;
  if (!c->psum) {
    int rc = dalloc(c, &c->psum, 7 * (size_t)n);
    if (rc) return rc;
  }
  HIPCHK(c, c3d::launch_plane_sums(c->g, c->U, c->V, c->W, c->psum, c->stream));
  m.assign(7 * (size_t)n, 0.0);
  HIPCHK(c, hipMemcpyAsync(m.data(), c->psum, m.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  int rc = sync(c);
  if (rc) return rc;
  const double nuv = (double)((c->nxg - 3) * (c->nzg - 3)), nw = (double)((c->nxg - 3) * (c->nz - 2));
  for (int q = 0; q < 7; q++)
    for (int j = 0; j < n; j++) m[(size_t)q * n + j] /= (q == 2 || q == 5) ? nw : nuv;
  return 0;
}

int c3d_monitor(c3d_ctx* c, double* out) {
  if (!c || !out) return C3D_EINVAL;
  if (!c->mon) {
    int rc = dalloc(c, &c->mon, (size_t)c3d::monitor_part_size() + 8);
    if (rc) return rc;
  }
  double* res = c->mon + c3d::monitor_part_size();
  HIPCHK(c, c3d::launch_monitor(c->g, c->U, c->V, c->W, c->mon, res, c->stream));
  double r[7];
  HIPCHK(c, hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  // src/monitor.f90:27-39, 57-59: maxima, and sums over nxm*nym*nzm with
  // nxm = nx - 1 etc. (src/initialization.f90:79-81), as written
  const double vol = (double)(c->nx - 1) * (double)(c->ny - 1) * (double)(c->nz - 1);
  out[0] = r[3];
  out[1] = r[4];
  out[2] = r[5];
  out[3] = r[0] / vol;
  out[4] = r[1] / vol;
  out[5] = r[2] / vol;
  out[6] = r[6];
  return C3D_OK;
}

// ---- allocation-free field summaries ----------------------------------------
// Device scratch of c3d_monitor_noalloc: the 256 x 7 per-workgroup partials
// and the 7 reduced sums of the monitor kernels (the same count whatever the
// grid).  Each device holds its own copy for the life of the process; its
// address is looked up once per device and kept in kScratchPtr.
namespace {
constexpr size_t kScratchDoubles = 2048;
static_assert((size_t)c3d::kMonitorPartDoubles + 8 <= kScratchDoubles, "monitor scratch");
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_ptr[kMaxDevices] = {nullptr};

// Address of this device's scratch (hipGetSymbolAddress on the first call
// on each device, with that device current), or null with *err set.
double* device_scratch(hipError_t* err) {
  int dev = -1;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  return scratch_ptr[dev];
}
}  // namespace

int c3d_monitor_noalloc(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                        double* out, void* stream) {
  if (nx < 3 || ny < 3 || nz < 3 || nx > 1000000 || ny > 1000000 || nz > 1000000) return C3D_EINVAL;
  if (!U || !V || !W || !out) return C3D_EINVAL;
  // the partials are a fixed 256 x 7 for every grid; this holds for the
  // kernels as built (the static_assert above), checked here as a limit
  if ((size_t)c3d::monitor_part_size() + 8 > kScratchDoubles) return C3D_EINVAL;
  hipError_t e = hipSuccess;
  double* part = device_scratch(&e);
  if (!part) {
    (void)hipGetLastError();
    return C3D_EHIP;
  }
  DevGrid g{};
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nxg = nx + 1; g.nyg = ny + 1; g.nzg = nz + 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* r7 = part + c3d::monitor_part_size();
  if (c3d::launch_monitor(g, U, V, W, part, r7, st) != hipSuccess) return C3D_EHIP;
  // src/monitor.f90:57-59: sums over nxm*nym*nzm, nxm = nx - 1 etc.
  const double vol = (double)(nx - 1) * (double)(ny - 1) * (double)(nz - 1);
  if (c3d::launch_monitor_finish(r7, vol, out, st) != hipSuccess) return C3D_EHIP;
  return C3D_OK;
}

int c3d_plane_statistics(c3d_ctx* c, double* means) {
  if (!c || !means) return C3D_EINVAL;
  std::vector<double> m;
  int rc = plane_means(c, m);
  if (rc) return rc;
  std::memcpy(means, m.data(), m.size() * sizeof(double));
  return C3D_OK;
}

int c3d_wall_shear(c3d_ctx* c, double* utau, double* wtau, double* Umean, double* Wmean) {
  if (!c || !utau || !wtau) return C3D_EINVAL;
  const int n = c->nyg;
  std::vector<double> m;
  int rc = plane_means(c, m);
  if (rc) return rc;
  const double* um = m.data();
  const double* wm = m.data() + 2 * (size_t)n;
  // wall gradients and friction velocities (src/statistics.f90:114-127); yg(j) is c->yg[j-1]
  const std::vector<double>& yg = c->yg;
  const double dub = (um[1] - um[0]) / (yg[1] - yg[0]);
  const double dut = (um[n - 1] - um[n - 2]) / (yg[n - 1] - yg[n - 2]);
  const double dwb = (wm[1] - wm[0]) / (yg[1] - yg[0]);
  const double dwt = (wm[n - 1] - wm[n - 2]) / (yg[n - 1] - yg[n - 2]);
  const double nu = c->desc.nu, uv = 0.0, vw = 0.0;  // wall Reynolds stresses are zero
  *utau = std::sqrt((uv - uv - nu * dut + nu * dub) / c->Ly);
  *wtau = std::sqrt((vw - vw - nu * dwt + nu * dwb) / c->Ly);
  if (Umean) std::memcpy(Umean, um, (size_t)n * sizeof(double));
  if (Wmean) std::memcpy(Wmean, wm, (size_t)n * sizeof(double));
  return C3D_OK;
}

int c3d_add_to_u(c3d_ctx* c, double d) {
  if (!c) return C3D_EINVAL;
  HIPCHK(c, c3d::launch_add_u(c->g, c->U, d, c->stream));
  return sync(c);
}

int c3d_time_step_rk3(c3d_ctx* c, double dt, int x_mass_cte, double q0) {
  if (!c || !(dt > 0.0)) return C3D_EINVAL;
  int rc = ensure_rk(c);
  if (!rc && hipMemsetAsync(c->nanflag, 0, sizeof(unsigned), c->stream) != hipSuccess)
    rc = fail(c, C3D_EHIP, "non-finite flag reset");
  const bool exact = c->rhs_exact;
  if (!rc && exact) rc = save_state(c);  // src/time_integration.f90:91-94 (fused: kept in place)
  const double t0 = c->t;
  static const double rk_t[3] = {8.0 / 15.0, 2.0 / 3.0, 1.0};  // src/initialization.f90:408-410
  for (int st = 1; st <= 3 && !rc; st++) {
    if (exact) {
      rc = rhs_launch(c, st);                                   // :101-103
      if (!rc) rc = rk_update_launch(c, st, dt);                // :105-107
    } else {
      rc = rk_stage_fast(c, st, dt);                            // :101-107 in one pass
    }
    c->t = t0 + rk_t[st - 1] * dt;                              // :108
    if (!rc && c->ib && c->ib->moving) rc = ib_move_to(c, c->t);  // :110-115
    if (!rc && c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream) != hipSuccess)
      rc = fail(c, C3D_EHIP, "boundary conditions");            // :117
    if (!rc) rc = save_interim(c);
    if (!rc) rc = projection_launch(c, dt, st, nullptr);        // :118
    if (!rc && c->ib) {                                         // :119-122
      if (c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream) != hipSuccess)
        rc = fail(c, C3D_EHIP, "boundary conditions");
      if (!rc) rc = ib_projection_launch(c, dt, st);
      if (rc == C3D_ENOCONV) rc = 0;  // the reference warns and continues
    }
    if (!rc && c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream) != hipSuccess)
      rc = fail(c, C3D_EHIP, "boundary conditions");            // :123
  }
  if (!rc && x_mass_cte == 1) {  // :187-193
    const double prev = c->dPdx;
    double ub = 0.0;
    rc = mass_flow(c, c->U, &ub);
    if (!rc) {
      const double d = q0 - ub;
      if (c3d::launch_add_u(c->g, c->U, d, c->stream) != hipSuccess ||
          c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream) != hipSuccess)
        rc = fail(c, C3D_EHIP, "mass-flow correction");
      c->dPdx = 0.5 * (prev + d / dt);
    }
  }
  // NaN stop of compute_statistics (src/statistics.f90:142-144), checked at
  // every step: the velocity corrections flag any non-finite value they write
  unsigned bad = 0;
  if (!rc && hipMemcpyAsync(&bad, c->nanflag, sizeof(bad), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = fail(c, C3D_EHIP, "non-finite flag read");
  if (!rc) rc = sync(c);
  if (!rc && bad) rc = fail(c, C3D_ENAN, "non-finite velocity after the time step at t = %.6e", c->t);
  return rc;
}

int c3d_set_rhs_exact(c3d_ctx* c, int on) {
  if (!c) return C3D_EINVAL;
  c->rhs_exact = on != 0;
  return C3D_OK;
}

int c3d_set_refinement(c3d_ctx* c, int nref) {
  if (!c) return C3D_EINVAL;
  if (nref < 0 || nref > kMaxRef) return fail(c, C3D_EINVAL, "nref must be 0..%d", kMaxRef);
  if (int rc = refine_buffers(c, nref)) return fail(c, rc, "refinement buffers: allocation failed");
  c->nref = nref;
  return C3D_OK;
}

int c3d_refinement_residuals(c3d_ctx* c, double* rmax, int n) {
  if (!c || !rmax || n < 0) return C3D_EINVAL;
  const int m = n < c->nref ? n : c->nref;
  std::vector<unsigned long long> bits((size_t)kMaxRef * kResSlots, 0ull);
  HIPCHK(c, hipMemcpyAsync(bits.data(), c->red + kRedRefine, bits.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  for (int s = 0; s < m; s++) {  // non-negative doubles order as their bit patterns
    unsigned long long mx = 0;
    for (int q = 0; q < kResSlots; q++) mx = std::max(mx, bits[(size_t)s * kResSlots + q]);
    std::memcpy(&rmax[s], &mx, sizeof(double));
  }
  return m;
}

int c3d_stage_count(void) { return kNumStages; }
const char* c3d_stage_name(int i) { return (i >= 0 && i < kNumStages) ? kStageNames[i] : ""; }

int c3d_time_projection(c3d_ctx* c, int iters, double dt, double* total_ms, double* kernel_ms,
                        int nk) {
  if (!c || iters <= 0 || !(dt != 0.0)) return C3D_EINVAL;
  std::vector<hipEvent_t> ev((size_t)iters * (kNumStages + 1));
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  int rc = 0;
  for (int it = 0; it < iters && !rc; it++)
    rc = projection_launch(c, dt, 2, &ev[(size_t)it * (kNumStages + 1)]);
  if (!rc) rc = sync(c);
  if (!rc) {
    std::vector<double> acc(kNumStages, 0.0);
    double tot = 0.0;
    for (int it = 0; it < iters && !rc; it++) {
      hipEvent_t* e = &ev[(size_t)it * (kNumStages + 1)];
      for (int s = 0; s < kNumStages; s++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e[s], e[s + 1]) != hipSuccess) {
          rc = fail(c, C3D_EHIP, "hipEventElapsedTime failed");
          break;
        }
        acc[s] += ms;
      }
      float ms = 0.f;
      if (!rc && hipEventElapsedTime(&ms, e[0], e[kNumStages]) == hipSuccess) tot += ms;
    }
    if (!rc) {
      if (total_ms) *total_ms = tot / iters;
      if (kernel_ms)
        for (int s = 0; s < nk && s < kNumStages; s++) kernel_ms[s] = acc[s] / iters;
    }
  }
  // second pass: the same number of projections between two events only, so
  // the per-stage event records add nothing to the total
  if (!rc && total_ms) {
    hipEvent_t& t0 = ev[0];
    hipEvent_t& t1 = ev[1];
    if (hipEventRecord(t0, c->stream) != hipSuccess) rc = fail(c, C3D_EHIP, "hipEventRecord failed");
    for (int it = 0; it < iters && !rc; it++) rc = projection_launch(c, dt, 2, nullptr);
    float ms = 0.f;
    if (!rc && (hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
      rc = fail(c, C3D_EHIP, "projection timing failed");
    if (!rc) *total_ms = ms / iters;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

// ----------------------------------------------------------------------------
// Immersed boundary (src/immersed_boundary_operators.f90, src/heaviside.f90,
// src/projection.f90:48-186, src/immersed_boundary_geometry.f90:400-442)
// ----------------------------------------------------------------------------

// Poisson solve of div(Ureg, Vreg, Wreg) into ib->phi = solution / dt, with
// the rk_step 1 pressure side effect (src/poisson.f90:91-99).
static int ib_poisson_reg(c3d_ctx* c, double dt, int rk_step) {
  IBState* ib = c->ib;
  const bool rk1 = rk_step == 1;
  HIPCHK(c, c3d::launch_fwd_x(c->g, true, ib->Ureg, ib->Vreg, ib->Wreg, nullptr, c->S, c->stream));
  int rc = spectral_core(c);
  if (rc) return rc;
  HIPCHK(c, c3d::launch_inv_x(c->g, 1, c->S, ib->phi, c->P, dt, rk1, c->stream));
  if (rk1) HIPCHK(c, c3d::launch_p_ghosts(c->g, c->P, c->stream));
  return 0;
}

// U_reg = R f (src/projection.f90:66-73): aux, three spreads, boundary conditions.
static int ib_spread_force(c3d_ctx* c, const double* f, double dt) {
  IBState* ib = c->ib;
  const int nb = ib->m.nb;
  HIPCHK(c, c3d::ib::launch_aux(ib->m, f, dt, c->desc.nu, ib->aux, c->stream));
  HIPCHK(c, c3d::ib::launch_spread3(ib->m, ib->sp, ib->aux, ib->Ureg, ib->Vreg, ib->Wreg, c->stream));
  HIPCHK(c, c3d::launch_apply_bc(c->g, ib->Ureg, ib->Vreg, ib->Wreg, c->stream));
  return 0;
}

// schur (src/projection.f90:98-144) on device vectors.
static int ib_schur_launch(c3d_ctx* c, const double* f, double* Sf, double dt, int rk_step) {
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb;
  int rc = ib_spread_force(c, f, dt);
  if (!rc) rc = ib_poisson_reg(c, dt, rk_step);
  if (rc) return rc;
  HIPCHK(c, c3d::ib::launch_schur_rows(c->g, ib->m, ib->phi, dt, ib->Ureg, ib->Vreg, ib->Wreg,
                                       ib->E1nH, f, Sf, c->stream));
  HIPCHK(c, c3d::ib::launch_remove_mean(ib->m, Sf + 3 * nb, ib->E1nHc, f + 3 * nb, ib->mpart, c->stream));
  return 0;
}

// One BiCGSTAB iteration (src/projection.f90:160-182) on the stream, ending
// with the squared residual norm stored to the pinned host word ib->herr.
static int ib_bicg_iteration(c3d_ctx* c, double* x, double dt, int rk_step) {
  IBState* ib = c->ib;
  namespace B = c3d::ib;
  const size_t n = 4 * (size_t)ib->m.nb;
  double *r = ib->vec[1], *rhat = ib->vec[2], *p = ib->vec[3], *v = ib->vec[4], *h = ib->vec[5],
         *sv = ib->vec[6], *t = ib->vec[7];
  int rc = 0;
  // partial slot 0 holds rhat.r on entry (the initial dot, then k_bicg_xr)
  HIPCHK(c, B::launch_bicg_p(r, v, ib->sc, ib->part, p, n, c->stream));
  if ((rc = ib_schur_launch(c, p, v, dt, rk_step))) return rc;
  HIPCHK(c, B::launch_dot(rhat, v, nullptr, nullptr, n, ib->part, 1, c->stream));
  HIPCHK(c, B::launch_bicg_hs(x, p, r, v, ib->sc, ib->part, h, sv, n, c->stream));
  if ((rc = ib_schur_launch(c, sv, t, dt, rk_step))) return rc;
  HIPCHK(c, B::launch_dot(t, sv, t, t, n, ib->part, 2, c->stream));
  HIPCHK(c, B::launch_bicg_xr(h, sv, t, rhat, ib->sc, ib->part, x, r, n, c->stream));
  // the scalar kernel stores ||r||^2 straight into the pinned host word, so a
  // captured iteration holds kernel nodes only (no device-to-host copy node)
  HIPCHK(c, B::launch_bicg_scalar(4, ib->sc, ib->part, ib->herr, c->stream));
  return 0;
}

// The captured iteration for (x, dt, rk_step == 1), captured on first use.
static int ib_iteration_graph(c3d_ctx* c, double* x, double dt, int rk_step, hipGraphExec_t* out) {
  IBState* ib = c->ib;
  const bool rk1 = rk_step == 1;
  for (auto& gr : ib->graphs)
    if (gr.x == x && gr.dt == dt && gr.rk1 == rk1) {
      *out = gr.exec;
      return 0;
    }
  if (ib->graphs.size() >= 8) {  // bounded cache: start over
    for (auto& gr : ib->graphs) (void)hipGraphExecDestroy(gr.exec);
    ib->graphs.clear();
  }
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = ib_bicg_iteration(c, x, dt, rk_step);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(c, C3D_EHIP, "capture of the BiCGSTAB iteration: %s", hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess)
    return fail(c, C3D_EHIP, "instantiation of the BiCGSTAB iteration: %s", hipGetErrorString(ei));
  ib->graphs.push_back({x, dt, rk1, exec});
  *out = exec;
  return 0;
}

// bicgstab (src/projection.f90:146-186) on device vectors x (in/out), b.
// r_ready: vec[1] already holds b - S x (the learned start computes it).
static int ib_bicgstab_launch(c3d_ctx* c, double* x, const double* b, double tol, int maxit,
                              double dt, int rk_step, int* iters_out, double* err_out,
                              bool r_ready = false) {
  IBState* ib = c->ib;
  namespace B = c3d::ib;
  const size_t n = 4 * (size_t)ib->m.nb;
  double *r = ib->vec[1], *rhat = ib->vec[2], *p = ib->vec[3], *v = ib->vec[4];
  int rc = 0;
  // coherent (fine-grained) host memory: the kernel's store is visible to the
  // host once the stream is synchronised
  if (!ib->herr &&
      hipHostMalloc((void**)&ib->herr, sizeof(double), hipHostMallocCoherent) != hipSuccess) {
    ib->herr = nullptr;
    (void)hipGetLastError();
    return fail(c, C3D_ENOMEM, "pinned residual word");
  }
  if (!r_ready && (rc = ib_schur_launch(c, x, r, dt, rk_step))) return rc;
  HIPCHK(c, B::launch_bicg_init(r_ready ? nullptr : b, r, rhat, p, v, n, c->stream));
  HIPCHK(c, B::launch_bicg_scalar(0, ib->sc, ib->part, nullptr, c->stream));
  HIPCHK(c, B::launch_dot(rhat, r, nullptr, nullptr, n, ib->part, 0, c->stream));
  hipGraphExec_t exec = nullptr;
  if (ib->use_graphs && maxit > 0 && (rc = ib_iteration_graph(c, x, dt, rk_step, &exec))) return rc;
  double error = 1.0;
  const double eps = tol * tol;
  int iter = 0;
  while (iter < maxit && error >= eps) {
    if (exec)
      HIPCHK(c, hipGraphLaunch(exec, c->stream));
    else if ((rc = ib_bicg_iteration(c, x, dt, rk_step)))
      return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    error = *ib->herr;
    iter++;
  }
  ib->last_iters = iter;
  ib->accum_iters += iter;
  ib->last_err = error;
  if (iters_out) *iters_out = iter;
  if (err_out) *err_out = error;
  if (iter >= maxit && error >= eps)
    return fail(c, C3D_ENOCONV, "BiCGSTAB reached %d iterations with ||r|| = %.3e > %.3e", iter,
                std::sqrt(error), tol);
  return 0;
}

// ---- history-projection start (guess_mode 4) ----
// The Schur operator depends only on dt, nu and the geometry, so for a static
// body every converged solve (f_i, b_i) of one dt satisfies S f_i = b_i to the
// solver tolerance (ib_move_to clears the pairs when the body moves).  The start for a new right-hand side b is the force combination
// whose rhs combination is closest to b in the 2-norm.
static int ib_hist_reserve(c3d_ctx* c) {
  IBState* ib = c->ib;
  if (ib->hF[0]) return 0;
  const size_t n4 = 4 * (size_t)ib->m.nb;
  constexpr int H = c3d::ib::kHist;
  int rc = 0;
  for (int j = 0; j < H && !rc; j++)
    if ((rc = dalloc(c, &ib->hF[j], n4)) || (rc = dalloc(c, &ib->hB[j], n4)) ||
        (rc = dalloc(c, &ib->hQ[j], n4)) || (rc = dalloc(c, &ib->hX[j], n4)))
      break;
  if (!rc) rc = dalloc(c, &ib->hpart, (size_t)(H + 1) * c3d::ib::kDotBlocks);
  if (!rc) rc = dalloc(c, &ib->hcoef, (size_t)4 * (H + 1));
  if (rc) ib->hF[0] = nullptr;
  return rc;
}

static void ib_hist_check_dt(IBState* ib, double dt) {
  if (ib->hdt != dt) {
    ib->hcount = 0;
    ib->hhead = 0;
    ib->hdt = dt;
  }
}

// Appends the converged pair (fb, rhs_ib) of a solve with step dt.
static int ib_hist_push(c3d_ctx* c, double dt) {
  IBState* ib = c->ib;
  const size_t bytes = 4 * (size_t)ib->m.nb * 8;
  ib_hist_check_dt(ib, dt);
  HIPCHK(c, hipMemcpyAsync(ib->hF[ib->hhead], ib->fb, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ib->hB[ib->hhead], ib->rhs_ib, bytes, hipMemcpyDeviceToDevice, c->stream));
  ib->hhead = (ib->hhead + 1) % c3d::ib::kHist;
  if (ib->hcount < c3d::ib::kHist) ib->hcount++;
  return 0;
}

// fb = sum_j <Q_j, rhs_ib> X_j over the basis of the stored pairs (oldest
// first, Gram-Schmidt with a second pass); fb is left as it is (the previous
// force) while no pair of this dt is stored.
static int ib_hist_start(c3d_ctx* c, double dt) {
  IBState* ib = c->ib;
  namespace B = c3d::ib;
  constexpr int H = B::kHist;
  ib_hist_check_dt(ib, dt);
  const int cnt = ib->hcount;
  if (cnt == 0) return 0;
  const size_t n4 = 4 * (size_t)ib->m.nb;
  const size_t bytes = n4 * 8;
  double* ca = ib->hcoef;                // first-pass coefficients; [i] = raw norm^2
  double* cb = ib->hcoef + (H + 1);      // second-pass coefficients
  double* cn = ib->hcoef + 2 * (H + 1);  // squared norm after orthogonalisation
  double* cf = ib->hcoef + 3 * (H + 1);  // coefficients of the start
  for (int i = 0; i < cnt; i++) {
    const int src = (ib->hhead - cnt + i + H) % H;
    HIPCHK(c, hipMemcpyAsync(ib->hQ[i], ib->hB[src], bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(ib->hX[i], ib->hF[src], bytes, hipMemcpyDeviceToDevice, c->stream));
    B::VecSet A{}, X{};
    for (int j = 0; j < i; j++) {
      A.v[j] = ib->hQ[j];
      X.v[j] = ib->hX[j];
    }
    A.v[i] = ib->hQ[i];
    A.k = i + 1;  // projections on Q_0..Q_{i-1} and the raw norm
    HIPCHK(c, B::launch_mdot(A, ib->hQ[i], n4, ib->hpart, c->stream));
    HIPCHK(c, B::launch_mreduce(ib->hpart, i + 1, ca, c->stream));
    if (i == 0) {
      HIPCHK(c, B::launch_hscale(ib->hQ[0], ib->hX[0], ca, ca, n4, c->stream));
      continue;
    }
    A.k = i;
    X.k = i;
    HIPCHK(c, B::launch_combine(A, X, ca, -1.0, false, ib->hQ[i], ib->hX[i], n4, c->stream));
    HIPCHK(c, B::launch_mdot(A, ib->hQ[i], n4, ib->hpart, c->stream));
    HIPCHK(c, B::launch_mreduce(ib->hpart, i, cb, c->stream));
    HIPCHK(c, B::launch_combine(A, X, cb, -1.0, false, ib->hQ[i], ib->hX[i], n4, c->stream));
    B::VecSet Nq{};
    Nq.v[0] = ib->hQ[i];
    Nq.k = 1;
    HIPCHK(c, B::launch_mdot(Nq, ib->hQ[i], n4, ib->hpart, c->stream));
    HIPCHK(c, B::launch_mreduce(ib->hpart, 1, cn, c->stream));
    HIPCHK(c, B::launch_hscale(ib->hQ[i], ib->hX[i], cn, ca + i, n4, c->stream));
  }
  B::VecSet A{}, X{};
  for (int j = 0; j < cnt; j++) {
    A.v[j] = ib->hQ[j];
    X.v[j] = ib->hX[j];
  }
  A.k = X.k = cnt;
  HIPCHK(c, B::launch_mdot(A, ib->rhs_ib, n4, ib->hpart, c->stream));
  HIPCHK(c, B::launch_mreduce(ib->hpart, cnt, cf, c->stream));
  HIPCHK(c, B::launch_combine(A, X, cf, 1.0, true, nullptr, ib->fb, n4, c->stream));
  return 0;
}

// compute_IB_projection (src/projection.f90:48-96).
static int ib_projection_launch(c3d_ctx* c, double dt, int rk_step) {
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb;
  HIPCHK(c, c3d::ib::launch_rhs_u(c->g, ib->m, c->U, c->V, c->W, ib->rhs_ib, c->stream));  // :51-52
  HIPCHK(c, c3d::ib::launch_interp_c(c->g, ib->m, c->Pint, -1.0, ib->rhs_ib + 3 * nb,
                                     c->stream));                                        // :53-54
  HIPCHK(c, c3d::ib::launch_remove_mean(ib->m, ib->rhs_ib + 3 * nb, nullptr, nullptr,
                                        ib->mpart, c->stream));                                     // :57
  if (ib->guess == 1) HIPCHK(c, hipMemsetAsync(ib->fb, 0, 4 * nb * 8, c->stream));
  if (ib->guess == 2) ib->guess = 0;
  // mode 5 starts from the converged force of the same RK3 stage of the
  // previous step (the stage forces of consecutive steps differ by O(dt),
  // those of consecutive stages by O(1))
  const bool by_stage = ib->guess == 5 && rk_step >= 1 && rk_step <= 3;
  if (by_stage) {
    if (!ib->fstage[0]) {
      int rc0 = 0;
      for (int q = 0; q < 3 && !rc0; q++) rc0 = dalloc(c, &ib->fstage[q], 4 * nb);
      if (rc0) {
        ib->fstage[0] = nullptr;
        return rc0;
      }
    }
    if (ib->fstage_ok[rk_step - 1])
      HIPCHK(c, hipMemcpyAsync(ib->fb, ib->fstage[rk_step - 1], 4 * nb * 8,
                               hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(ib->fprev, ib->fb, 4 * nb * 8, hipMemcpyDeviceToDevice, c->stream));
  ib->last_dt = dt;
  bool r_ready = false;
  if (ib->guess == 3) {  // learned start: fb = fprev + a d, d from the network
    int rc0 = mlp_guess(c, dt, rk_step);
    if (rc0) return rc0;
    r_ready = true;
  }
  if (ib->guess == 4) {  // least-squares start over earlier solves of this dt
    int rc0 = ib_hist_reserve(c);
    if (!rc0) rc0 = ib_hist_start(c, dt);
    if (rc0) return rc0;
  }
  int rc = ib_bicgstab_launch(c, ib->fb, ib->rhs_ib, ib->tol, ib->maxit, dt, rk_step, nullptr,
                              nullptr, r_ready);                                         // :60
  if (rc && rc != C3D_ENOCONV) return rc;
  const int conv = rc;
  if (!conv && ib->guess == 4 && (rc = ib_hist_push(c, dt))) return rc;
  if (by_stage) {
    HIPCHK(c, hipMemcpyAsync(ib->fstage[rk_step - 1], ib->fb, 4 * nb * 8, hipMemcpyDeviceToDevice,
                             c->stream));
    ib->fstage_ok[rk_step - 1] = true;
  }
  if ((rc = ib_spread_force(c, ib->fb, dt))) return rc;                                  // :66-73
  if ((rc = ib_poisson_reg(c, dt, rk_step))) return rc;                                  // :76-80
  HIPCHK(c, c3d::ib::launch_ib_correct(c->g, ib->Uint, ib->Vint, ib->Wint, ib->Ureg, ib->Vreg,
                                       ib->Wreg, c->Pint, ib->phi, dt, c->U, c->V, c->W,
                                       c->stream));                                      // :83-92
  HIPCHK(c, c3d::launch_apply_bc(c->g, c->U, c->V, c->W, c->stream));                  // :94
  return conv;  // C3D_ENOCONV keeps the message set by the solve
}

static int need_ib(c3d_ctx* c) {
  if (!c) return C3D_EINVAL;
  if (!c->ib) return fail(c, C3D_EINVAL, "no immersed body set up (c3d_ib_setup)");
  return 0;
}

}  // extern "C"

namespace {

// Marker, stencil and resident field storage for nb markers (one body set-up
// per context).  On failure c->ib keeps the partial state for ib_discard.
int ib_create(c3d_ctx* c, int nb, int nzb, int nxb, int nxb1) {
  IBState* ib = new IBState();
  c->ib = ib;
  {  // C3D_IB_GRAPHS=0: launch each BiCGSTAB iteration kernel by kernel
    const char* e = getenv("C3D_IB_GRAPHS");
    ib->use_graphs = !(e && atoi(e) == 0);
  }
  namespace B = c3d::ib;
  B::Markers& m = ib->m;
  m.nb = nb; m.nzb = nzb; m.nxb = nxb; m.nxb1 = nxb1;
  m.A = c->dx * c->dymin * c->dz;
  m.factor = 1.0 / (c->dx * c->dymin * c->dz);
  const size_t n4 = 4 * (size_t)nb, n3 = 3 * (size_t)nb;
  int rc = 0;
  if ((rc = dalloc(c, &ib->xb, nb)) || (rc = dalloc(c, &ib->yb, nb)) ||
      (rc = dalloc(c, &ib->zb, nb)) || (rc = dalloc(c, &ib->normals, n3)) ||
      (rc = dalloc(c, &ib->sb, nb)) || (rc = dalloc(c, &ib->ub, n3)) ||
      (rc = dalloc(c, &ib->yref, nb)) || (rc = dalloc(c, &ib->flags, 8)))
    return rc;
  m.xb = ib->xb; m.yb = ib->yb; m.zb = ib->zb; m.normals = ib->normals; m.sb = ib->sb;
  m.ub = ib->ub; m.yref = ib->yref;
  for (int s = 0; s < 4; s++)
    if ((rc = dalloc(c, &m.st[s].idx, 15 * (size_t)nb)) || (rc = dalloc(c, &m.st[s].phi, 15 * (size_t)nb)) ||
        (rc = dalloc(c, &m.st[s].dd, 15 * (size_t)nb)))
      return rc;
  B::SetupGrid& sg = ib->sg;
  if ((rc = upload(c, &sg.x, c->x)) || (rc = upload(c, &sg.y, c->y)) || (rc = upload(c, &sg.z, c->z)))
    return rc;
  sg.dx = c->dx; sg.dz = c->dz; sg.dymin = c->dymin; sg.Lxp = c->Lxp; sg.Lzp = c->Lzp;
  // resident buffers
  if ((rc = dalloc(c, &ib->Ureg, c->nu)) || (rc = dalloc(c, &ib->Vreg, c->nv)) ||
      (rc = dalloc(c, &ib->Wreg, c->nw)) || (rc = dalloc(c, &ib->phi, c->nr)) ||
      (rc = dalloc(c, &ib->Hc_ext, c->nr)) || (rc = dalloc(c, &ib->Hu_int, c->nu)) ||
      (rc = dalloc(c, &ib->Hv_int, c->nv)) || (rc = dalloc(c, &ib->Hw_int, c->nw)) ||
      (rc = dalloc(c, &ib->Uint, c->nu)) || (rc = dalloc(c, &ib->Vint, c->nv)) ||
      (rc = dalloc(c, &ib->Wint, c->nw)) || (rc = dalloc(c, &ib->E1nHc, (size_t)nb)) ||
      (rc = dalloc(c, &ib->E1nH, n3)) || (rc = dalloc(c, &ib->aux, n3)) ||
      (rc = dalloc(c, &ib->rhs_ib, n4)) || (rc = dalloc(c, &ib->fb, n4)) ||
      (rc = dalloc(c, &ib->fin, n4)) || (rc = dalloc(c, &ib->fout, n4)) ||
      (rc = dalloc(c, &ib->fprev, n4)) ||
      (rc = dalloc(c, &ib->sc, 8)) || (rc = dalloc(c, &ib->part, 5 * (size_t)B::kDotBlocks)) ||
      (rc = dalloc(c, &ib->mpart, 128)))
    return rc;
  for (int v = 0; v < 8; v++)
    if ((rc = dalloc(c, &ib->vec[v], n4))) return rc;
  if (hipMemsetAsync(ib->fb, 0, n4 * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(ib->phi, 0, c->nr * 8, c->stream) != hipSuccess)
    return fail(c, C3D_EHIP, "IB buffer clear failed");
  return 0;
}

void ib_discard(c3d_ctx* c) {
  delete c->ib;  // device memory stays in c->allocs until c3d_destroy
  c->ib = nullptr;
}

// Stencils (setup_IB_operators, src/immersed_boundary_operators.f90:19-256)
// and the spreading maps built on the host from the stencil indices.
int ib_operators_host_map(c3d_ctx* c) {
  namespace B = c3d::ib;
  IBState* ib = c->ib;
  B::Markers& m = ib->m;
  const int nb = m.nb;
  int rc = 0;
  if (B::launch_setup(c->g, ib->sg, m, nullptr, c->stream) != hipSuccess)
    return fail(c, C3D_EHIP, "IB stencil set-up launch failed");
  std::vector<int> idx[4];
  for (int s = 0; s < 4; s++) {
    idx[s].resize(15 * (size_t)nb);
    if (hipMemcpyAsync(idx[s].data(), m.st[s].idx, idx[s].size() * 4, hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess)
      return fail(c, C3D_EHIP, "stencil copy-back failed");
  }
  if ((rc = sync(c))) return rc;
  // every node must be a canonical interior node inside the gradient's
  // ranges (the Schur kernels evaluate G at the nodes and skip the boundary
  // conditions there): x in [2, nx-1], z in [2, nz-1], y in [2, nyg-1]
  // (v: [2, ny-1])
  const int ylim[4] = {c->nyg - 1, c->ny - 1, c->nyg - 1, c->nyg - 1};
  for (int s = 0; s < 4; s++)
    for (int t = 0; t < 5; t++)
      for (int l = 0; l < nb; l++) {
        const int X = idx[s][(0 * 5 + t) * (size_t)nb + l], Y = idx[s][(1 * 5 + t) * (size_t)nb + l],
                  Z = idx[s][(2 * 5 + t) * (size_t)nb + l];
        if (X < 2 || X > c->nx - 1 || Z < 2 || Z > c->nz - 1)
          return fail(c, C3D_EINVAL, "stencil %d of marker %d: x/z index out of range", s, l);
        if (Y < 2 || Y > ylim[s])
          return fail(c, C3D_EINVAL, "marker %d: stencil reaches a wall ghost row", l);
      }
  // node -> entry maps of the three spreading operators (counting sort by
  // node offset; entries within a node stay in ascending marker order)
  const int n1[3] = {c->nx, c->nxg, c->nxg}, n2[3] = {c->nyg, c->ny, c->nyg};
  const size_t nf[3] = {c->nu, c->nv, c->nw};
  for (int q = 0; q < 3; q++) {
    const std::vector<int>& I = idx[q];
    std::vector<int> cnt(nf[q] + 1, 0), off((size_t)nb * 125);
    for (int l = 0; l < nb; l++)
      for (int p = 0; p < 125; p++) {
        const int i = p % 5, j = (p / 5) % 5, k = p / 25;
        const size_t o = c3d::fidx(I[(0 * 5 + i) * (size_t)nb + l], I[(1 * 5 + j) * (size_t)nb + l],
                                   I[(2 * 5 + k) * (size_t)nb + l], n1[q], n2[q]);
        off[(size_t)l * 125 + p] = (int)o;
        cnt[o + 1]++;
      }
    // CSR by node offset first (entries in ascending e = l*125 + p)
    std::vector<int> node, ncnt;
    for (size_t o = 0; o < nf[q]; o++)
      if (cnt[o + 1]) {
        node.push_back((int)o);
        ncnt.push_back(cnt[o + 1]);
      }
    for (size_t o = 0; o < nf[q]; o++) cnt[o + 1] += cnt[o];
    std::vector<int> csr((size_t)nb * 125);
    for (size_t e = 0; e < csr.size(); e++) csr[cnt[off[e]]++] = (int)e;
    // jagged diagonal order: ranks by descending count (stable in offset)
    const size_t nn = node.size();
    std::vector<int> start(nn), rank(nn);
    for (size_t r = 0, acc = 0; r < nn; r++) {
      start[r] = (int)acc;
      acc += ncnt[r];
      rank[r] = (int)r;
    }
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b2) { return ncnt[a] > ncnt[b2]; });
    const int maxc = nn ? ncnt[rank[0]] : 0;
    std::vector<int> slot(maxc + 1, 0), jnode(nn), jcnt(nn), qq(csr.size());
    {
      std::vector<int> per(maxc, 0);  // ranks with more than i entries
      for (size_t r = 0; r < nn; r++)
        for (int i = 0; i < ncnt[rank[r]]; i++) per[i]++;
      for (int i = 0; i < maxc; i++) slot[i + 1] = slot[i] + per[i];
    }
    for (size_t r = 0; r < nn; r++) {
      const int src = rank[r];
      jnode[r] = node[src];
      jcnt[r] = ncnt[src];
      for (int i = 0; i < ncnt[src]; i++) qq[(size_t)slot[i] + r] = csr[(size_t)start[src] + i];
    }
    ib->sp[q].nnodes = (int)nn;
    ib->sp[q].nentries = (int)qq.size();
    if ((rc = upload(c, &ib->sp[q].node, jnode)) || (rc = upload(c, &ib->sp[q].cnt, jcnt)) ||
        (rc = upload(c, &ib->sp[q].slot, slot)) || (rc = upload(c, &ib->sp[q].q, qq)) ||
        (rc = dalloc(c, &ib->sp[q].ml, qq.size())) || (rc = dalloc(c, &ib->sp[q].wfs, qq.size())))
      return rc;
    if (B::launch_spread_coeffs(m, ib->sp[q], q, c->stream) != hipSuccess)
      return fail(c, C3D_EHIP, "spreading coefficients");
  }
  return 0;
}

// Stencils and spreading maps built on the device (the moving-body path):
// no stencil index leaves the device; the node counts and two range flags
// come back in one 32-byte read.
int ib_operators_device_map(c3d_ctx* c) {
  namespace B = c3d::ib;
  IBState* ib = c->ib;
  B::Markers& m = ib->m;
  int hf[8] = {0};
  HIPCHK(c, hipMemsetAsync(ib->flags + 1, 0, 4, c->stream));
  HIPCHK(c, B::launch_setup(c->g, ib->sg, m, ib->flags + 1, c->stream));
  HIPCHK(c, hipMemcpyAsync(hf, ib->flags, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  if (hf[1])
    return fail(c, C3D_EINVAL, "marker %d: stencil leaves the interior rows (wall ghost row)",
                hf[1] - 1);
  const int n1[3] = {c->nx, c->nxg, c->nxg}, n2[3] = {c->nyg, c->ny, c->nyg};
  const size_t nf[3] = {c->nu, c->nv, c->nw};
  for (int q = 0; q < 3; q++) {
    HIPCHK(c, B::build_spread_map(m, q, n1[q], n2[q], nf[q], ib->map_work, ib->sp[q],
                                  ib->flags + 2 + q, c->stream));
    HIPCHK(c, B::launch_spread_coeffs(m, ib->sp[q], q, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(hf, ib->flags, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if ((rc = sync(c))) return rc;
  for (int q = 0; q < 3; q++) ib->sp[q].nnodes = hf[2 + q];
  return 0;
}

// compute_heaviside (src/heaviside.f90:20-54) for the current markers.  The
// spread targets are cleared first: the spreading kernels write only the
// support nodes, which move with the body.
int ib_heaviside(c3d_ctx* c) {
  namespace B = c3d::ib;
  IBState* ib = c->ib;
  B::Markers& m = ib->m;
  const int nb = m.nb;
  int rc = 0;
  if (hipMemsetAsync(ib->Ureg, 0, c->nu * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(ib->Vreg, 0, c->nv * 8, c->stream) != hipSuccess ||
      hipMemsetAsync(ib->Wreg, 0, c->nw * 8, c->stream) != hipSuccess)
    return fail(c, C3D_EHIP, "IB buffer clear failed");
  // compute_heaviside (src/heaviside.f90:20-54); the exterior face fields
  // are built in the U_interim buffers
  double *Hu = ib->Uint, *Hv = ib->Vint, *Hw = ib->Wint;
  double* Freg[3] = {ib->Ureg, ib->Vreg, ib->Wreg};
  bool ok = true;
  for (int q = 0; q < 3 && ok; q++)
    ok = B::launch_spread(m, ib->sp[q], q, m.normals + (size_t)q * nb, Freg[q], c->stream) == hipSuccess;
  ok = ok && c3d::launch_apply_bc(c->g, ib->Ureg, ib->Vreg, ib->Wreg, c->stream) == hipSuccess;
  ok = ok && c3d::launch_divergence(c->g, ib->Ureg, ib->Vreg, ib->Wreg, ib->Hc_ext, c->stream) == hipSuccess;
  ok = ok && B::launch_affine(ib->Hc_ext, -0.0, -1.0, ib->Hc_ext, c->nr, c->stream) == hipSuccess;
  ok = ok && c3d::launch_fwd_x(c->g, false, nullptr, nullptr, nullptr, ib->Hc_ext, c->S, c->stream) == hipSuccess;
  if (!ok) return fail(c, C3D_EHIP, "Heaviside set-up failed");
  if ((rc = spectral_core(c))) return rc;
  ok = c3d::launch_inv_x(c->g, 0, c->S, ib->Hc_ext, c->P, 1.0, false, c->stream) == hipSuccess;
  ok = ok && B::launch_affine(ib->Hc_ext, 1.0, -1.0, ib->Hc_ext, c->nr, c->stream) == hipSuccess;
  ok = ok && hipMemsetAsync(Hu, 0, c->nu * 8, c->stream) == hipSuccess &&
       hipMemsetAsync(Hv, 0, c->nv * 8, c->stream) == hipSuccess &&
       hipMemsetAsync(Hw, 0, c->nw * 8, c->stream) == hipSuccess;
  ok = ok && B::launch_h_interp(c->g, ib->Hc_ext, Hu, Hv, Hw, c->stream) == hipSuccess;
  ok = ok && c3d::launch_apply_bc(c->g, Hu, Hv, Hw, c->stream) == hipSuccess;
  ok = ok && B::launch_affine(Hu, 1.0, -1.0, ib->Hu_int, c->nu, c->stream) == hipSuccess &&
       B::launch_affine(Hv, 1.0, -1.0, ib->Hv_int, c->nv, c->stream) == hipSuccess &&
       B::launch_affine(Hw, 1.0, -1.0, ib->Hw_int, c->nw, c->stream) == hipSuccess;
  ok = ok && c3d::launch_apply_bc(c->g, ib->Hu_int, ib->Hv_int, ib->Hw_int, c->stream) == hipSuccess;
  ok = ok && B::launch_interp_c(c->g, m, ib->Hc_ext, 1.0, ib->E1nHc, c->stream) == hipSuccess;
  ok = ok && B::launch_interp_uvw(c->g, m, Hu, Hv, Hw, true, ib->E1nH, c->stream) == hipSuccess;
  if (!ok) return fail(c, C3D_EHIP, "Heaviside set-up failed");
  if ((rc = sync(c))) return rc;
  c->Hm = ib->Hu_int;
  return 0;
}

// setup_IB_geometry of the moving walls at time t
// (src/immersed_boundary_geometry.f90:176-390) into the marker arrays.
int ib_wave_geometry(c3d_ctx* c, double t) {
  namespace B = c3d::ib;
  IBState* ib = c->ib;
  const B::WaveBody& w = ib->wave;
  B::WaveTime wt{};
  wt.t = t;
  wt.ct = std::cos(w.p2 * t);
  wt.st = std::sin(w.p2 * t);
  wt.amp = w.p1;
  wt.damp = 0.0;
  if (w.body == 2 && t < ib->ramp && ib->ramp > 0.0) {  // :274-281
    const double tau = t / ib->ramp;
    wt.amp = w.p1 * 0.5 * (1.0 - std::cos(w.pi * tau));
    wt.damp = w.p1 * 0.5 * (w.pi / ib->ramp) * std::sin(w.pi * tau);
  }
  int bad = 0;
  HIPCHK(c, hipMemsetAsync(ib->flags, 0, 4, c->stream));
  HIPCHK(c, B::launch_wave_geometry(w, wt, ib->xb, ib->yb, ib->zb, ib->normals, ib->sb, ib->ub,
                                    ib->yref, ib->flags, c->stream));
  HIPCHK(c, hipMemcpyAsync(&bad, ib->flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  if (bad)  // :199-215, :305-321
    return fail(c, C3D_EINVAL, "marker %d at t = %g: body points support exceeds grid dimensions",
                bad - 1, t);
  return 0;
}

// Geometry, operators and Heaviside at time t; the cached iteration graphs
// (launch shapes follow the node counts) and the history start (the Schur
// operator changed) are dropped.
int ib_move_to(c3d_ctx* c, double t) {
  IBState* ib = c->ib;
  int rc = ib_wave_geometry(c, t);
  if (!rc) rc = ib_operators_device_map(c);
  if (!rc) rc = ib_heaviside(c);
  if (rc) return rc;
  for (auto& gr : ib->graphs) (void)hipGraphExecDestroy(gr.exec);
  ib->graphs.clear();
  ib->hcount = 0;
  ib->hhead = 0;
  ib->t_geom = t;
  return 0;
}

}  // namespace

extern "C" {

int c3d_ib_setup_ref(c3d_ctx* c, int nb, const double* xb, const double* yb, const double* zb,
                     const double* normals, const double* sb, const double* ub, const int* yref,
                     int nzb, int nxb, int nxb1) {
  if (!c || !xb || !yb || !zb || !normals || !sb || !ub) return C3D_EINVAL;
  if (c->ib) return fail(c, C3D_EINVAL, "an immersed body is already set up on this context");
  if (nb <= 0 || nzb <= 0 || nxb <= 0 || (long long)nzb * nxb != nb || nxb1 <= 0 || nxb1 > nxb)
    return fail(c, C3D_EINVAL, "marker counts: nb = nzb*nxb and 0 < nxb1 <= nxb required");
  if ((long long)nb * 125 > 0x7fffffff) return fail(c, C3D_EINVAL, "too many markers");
  for (int l = 0; l < nb; l++) {
    if (!(xb[l] >= 0.0 && xb[l] < c->Lxp && zb[l] >= 0.0 && zb[l] < c->Lzp && yb[l] > c->y[0] &&
          yb[l] < c->y[c->ny - 1]))
      return fail(c, C3D_EINVAL, "marker %d outside the periodic box", l);
    if (yref && (yref[l] < 1 || yref[l] > c->ny))
      return fail(c, C3D_EINVAL, "marker %d: y_ref_index outside 1..ny", l);
  }
  int rc = ib_create(c, nb, nzb, nxb, nxb1);
  IBState* ib = c->ib;
  const size_t n3 = 3 * (size_t)nb;
  std::vector<int> yr(nb, 1);
  if (yref) yr.assign(yref, yref + nb);
  if (!rc &&
      (hipMemcpyAsync(ib->xb, xb, nb * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->yb, yb, nb * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->zb, zb, nb * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->normals, normals, n3 * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->sb, sb, nb * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->ub, ub, n3 * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(ib->yref, yr.data(), nb * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess))
    rc = fail(c, C3D_EHIP, "marker upload failed");
  if (!rc) rc = ib_operators_host_map(c);
  if (!rc) rc = ib_heaviside(c);
  if (rc) ib_discard(c);
  return rc;
}

int c3d_ib_setup(c3d_ctx* c, int nb, const double* xb, const double* yb, const double* zb,
                 const double* normals, const double* sb, const double* ub, int nzb, int nxb,
                 int nxb1) {
  return c3d_ib_setup_ref(c, nb, xb, yb, zb, normals, sb, ub, nullptr, nzb, nxb, nxb1);
}

int c3d_ib_setup_wave(c3d_ctx* c, int body, int nxb, int nzb, double p1, double p2, double p3,
                      double ramp, double t) {
  if (!c) return C3D_EINVAL;
  if (c->ib) return fail(c, C3D_EINVAL, "an immersed body is already set up on this context");
  if (body != 1 && body != 2) return fail(c, C3D_EINVAL, "body must be 1 (standing_wave) or 2 (traveling_wave_x)");
  if (nxb < 3 || nzb < 1) return fail(c, C3D_EINVAL, "nxb >= 3 and nzb >= 1 required");
  if ((long long)2 * nxb * nzb * 125 >= 0x7fffffff) return fail(c, C3D_EINVAL, "too many markers");
  if (body == 1 && p1 > c->desc.min_buffer_width)  // src/immersed_boundary_geometry.f90:177
    return fail(c, C3D_EINVAL, "IB amplitude is bigger than the minimum buffer width");
  if (body == 2 && !(p2 * p3 != 0.0))
    return fail(c, C3D_EINVAL, "traveling wave: body_param_2 * body_param_3 must be nonzero");
  const int nb = 2 * nxb * nzb;
  int rc = ib_create(c, nb, nzb, 2 * nxb, nxb);
  IBState* ib = c->ib;
  namespace B = c3d::ib;
  if (!rc) {
    ib->moving = true;
    // a move invalidates the captured iteration graphs; recapturing them every
    // stage costs more than replaying saves (47.4 against 46.0 ms per RK3 step
    // at nb = 16,384), so a moving body launches kernel by kernel unless
    // c3d_ib_set_graphs turns the graphs back on
    ib->use_graphs = false;
    ib->ramp = ramp;
    B::WaveBody& w = ib->wave;
    w.body = body; w.nxb = nxb; w.nzb = nzb; w.ny = c->ny;
    w.p1 = p1; w.p2 = p2; w.p3 = p3;
    w.Lxp = c->Lxp; w.Lzp = c->Lzp; w.Lyc = c->desc.Ly; w.pi = 4.0 * std::atan(1.0);
    w.ylo = c->y[2];
    w.yhi = c->y[c->ny - 3];
    // spreading maps at their capacity: nb*125 entries, as many nodes
    const size_t ne = (size_t)nb * 125;
    size_t nfmax = std::max(c->nu, std::max(c->nv, c->nw));
    for (int q = 0; q < 3 && !rc; q++) {
      B::Spread& sp = ib->sp[q];
      int *node = nullptr, *cnt = nullptr, *slot = nullptr, *qq = nullptr;
      if ((rc = dalloc(c, &node, ne)) || (rc = dalloc(c, &cnt, ne)) ||
          (rc = dalloc(c, &slot, ne + 1)) || (rc = dalloc(c, &qq, ne)) ||
          (rc = dalloc(c, &sp.ml, ne)) || (rc = dalloc(c, &sp.wfs, ne)))
        break;
      sp.node = node; sp.cnt = cnt; sp.slot = slot; sp.q = qq;
      sp.nentries = (int)ne;
    }
    if (!rc) {
      unsigned char* work = nullptr;
      rc = dalloc(c, &work, B::spread_map_work_bytes(nb, nfmax));
      ib->map_work = work;
    }
  }
  if (!rc) rc = ib_move_to(c, t);
  if (rc) ib_discard(c);
  return rc;
}

int c3d_ib_move(c3d_ctx* c, double t) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!c->ib->moving) return fail(c, C3D_EINVAL, "the immersed body does not move");
  return ib_move_to(c, t);
}

int c3d_ib_get_markers(c3d_ctx* c, double* xb, double* yb, double* zb, double* normals,
                       double* sb, double* ub, int* yref) {
  int rc = need_ib(c);
  if (rc) return rc;
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb;
  const double* src[6] = {ib->xb, ib->yb, ib->zb, ib->normals, ib->sb, ib->ub};
  double* dst[6] = {xb, yb, zb, normals, sb, ub};
  const size_t n[6] = {nb, nb, nb, 3 * nb, nb, 3 * nb};
  for (int a = 0; a < 6; a++)
    if (dst[a]) HIPCHK(c, hipMemcpyAsync(dst[a], src[a], n[a] * 8, hipMemcpyDeviceToHost, c->stream));
  if (yref) HIPCHK(c, hipMemcpyAsync(yref, ib->yref, nb * 4, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_set_time(c3d_ctx* c, double t) {
  if (!c) return C3D_EINVAL;
  c->t = t;
  return C3D_OK;
}

int c3d_get_time(const c3d_ctx* c, double* t) {
  if (!c || !t) return C3D_EINVAL;
  *t = c->t;
  return C3D_OK;
}

int c3d_ib_get_stencil(c3d_ctx* c, int s, int* xi, int* yi, int* zi, double* w, double* dxn) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (s < 0 || s > 3) return fail(c, C3D_EINVAL, "stencil must be 0..3");
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb, n = 15 * nb;
  std::vector<int> I(n);
  std::vector<double> ph(n), dd(n), nrm(3 * nb);
  HIPCHK(c, hipMemcpyAsync(I.data(), ib->m.st[s].idx, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ph.data(), ib->m.st[s].phi, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(dd.data(), ib->m.st[s].dd, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(nrm.data(), ib->m.normals, 3 * nb * 8, hipMemcpyDeviceToHost, c->stream));
  if ((rc = sync(c))) return rc;
  for (size_t l = 0; l < nb; l++)
    for (int p = 0; p < 125; p++) {
      const size_t a = (size_t)(p % 5) * nb + l, b = (size_t)(5 + (p / 5) % 5) * nb + l,
                   d = (size_t)(10 + p / 25) * nb + l, o = l * 125 + p;
      if (xi) xi[o] = I[a];
      if (yi) yi[o] = I[b];
      if (zi) zi[o] = I[d];
      if (w) w[o] = ((ib->m.A * ph[a]) * ph[b]) * ph[d];
      if (dxn) {
#pragma clang fp contract(off)
        dxn[o] = (nrm[l] * dd[a] + nrm[nb + l] * dd[b]) + nrm[2 * nb + l] * dd[d];
      }
    }
  return C3D_OK;
}

int c3d_ib_get_heaviside(c3d_ctx* c, double* E1nHc, double* E1nH, double* Hu_int, double* Hv_int,
                         double* Hw_int, double* Hc_ext) {
  int rc = need_ib(c);
  if (rc) return rc;
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb;
  if (E1nHc) HIPCHK(c, hipMemcpyAsync(E1nHc, ib->E1nHc, nb * 8, hipMemcpyDeviceToHost, c->stream));
  if (E1nH) HIPCHK(c, hipMemcpyAsync(E1nH, ib->E1nH, 3 * nb * 8, hipMemcpyDeviceToHost, c->stream));
  if (Hu_int) HIPCHK(c, hipMemcpyAsync(Hu_int, ib->Hu_int, c->nu * 8, hipMemcpyDeviceToHost, c->stream));
  if (Hv_int) HIPCHK(c, hipMemcpyAsync(Hv_int, ib->Hv_int, c->nv * 8, hipMemcpyDeviceToHost, c->stream));
  if (Hw_int) HIPCHK(c, hipMemcpyAsync(Hw_int, ib->Hw_int, c->nw * 8, hipMemcpyDeviceToHost, c->stream));
  if (Hc_ext) HIPCHK(c, hipMemcpyAsync(Hc_ext, ib->Hc_ext, c->nr * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_interpolate(c3d_ctx* c, int with_dxn, double* out) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!out) return C3D_EINVAL;
  IBState* ib = c->ib;
  HIPCHK(c, c3d::ib::launch_interp_uvw(c->g, ib->m, c->U, c->V, c->W, with_dxn != 0, ib->fout,
                                       c->stream));
  HIPCHK(c, hipMemcpyAsync(out, ib->fout, 3 * (size_t)ib->m.nb * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_interpolate_arrays(c3d_ctx* c, int with_dxn, const double* U, const double* V,
                              const double* W, double* out) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!U || !V || !W || !out) return C3D_EINVAL;
  IBState* ib = c->ib;
  if ((rc = stage_uvw(c, U, V, W))) return rc;
  HIPCHK(c, c3d::ib::launch_interp_uvw(c->g, ib->m, c->xU, c->xV, c->xW, with_dxn != 0, ib->fout,
                                       c->stream));
  HIPCHK(c, hipMemcpyAsync(out, ib->fout, 3 * (size_t)ib->m.nb * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_compute_heaviside(c3d_ctx* c) {
  int rc = need_ib(c);
  if (rc) return rc;
  // the markers, stencils and spreading maps are unchanged, so cached
  // iteration graphs stay valid
  return ib_heaviside(c);
}

int c3d_ib_interpolate_c(c3d_ctx* c, const double* p, double* out) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!p || !out) return C3D_EINVAL;
  IBState* ib = c->ib;
  HIPCHK(c, hipMemcpyAsync(ib->phi, p, c->nr * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, c3d::ib::launch_interp_c(c->g, ib->m, ib->phi, 1.0, ib->fout, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, ib->fout, (size_t)ib->m.nb * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_spread(c3d_ctx* c, int comp, const double* f, double* F) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (comp < 0 || comp > 2 || !f || !F) return C3D_EINVAL;
  IBState* ib = c->ib;
  const size_t nf[3] = {c->nu, c->nv, c->nw};
  double* T[3] = {ib->Ureg, ib->Vreg, ib->Wreg};
  HIPCHK(c, hipMemcpyAsync(ib->fin, f, (size_t)ib->m.nb * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(T[comp], 0, nf[comp] * 8, c->stream));
  HIPCHK(c, c3d::ib::launch_spread(ib->m, ib->sp[comp], comp, ib->fin, T[comp], c->stream));
  HIPCHK(c, hipMemcpyAsync(F, T[comp], nf[comp] * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_remove_mean(c3d_ctx* c, double* f) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!f) return C3D_EINVAL;
  IBState* ib = c->ib;
  const size_t nb = ib->m.nb;
  HIPCHK(c, hipMemcpyAsync(ib->fin, f, nb * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, c3d::ib::launch_remove_mean(ib->m, ib->fin, nullptr, nullptr, ib->mpart, c->stream));
  HIPCHK(c, hipMemcpyAsync(f, ib->fin, nb * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_schur(c3d_ctx* c, const double* f, double* Sf, double dt, int rk_step) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!f || !Sf || !(dt != 0.0)) return C3D_EINVAL;
  IBState* ib = c->ib;
  const size_t n4 = 4 * (size_t)ib->m.nb;
  HIPCHK(c, hipMemcpyAsync(ib->fin, f, n4 * 8, hipMemcpyHostToDevice, c->stream));
  if ((rc = ib_schur_launch(c, ib->fin, ib->fout, dt, rk_step))) return rc;
  HIPCHK(c, hipMemcpyAsync(Sf, ib->fout, n4 * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_ib_bicgstab(c3d_ctx* c, double* x, const double* b, double tol, int maxit, double dt,
                    int rk_step, int* iters, double* err) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!x || !b || maxit < 0 || !(dt != 0.0)) return C3D_EINVAL;
  IBState* ib = c->ib;
  const size_t n4 = 4 * (size_t)ib->m.nb;
  HIPCHK(c, hipMemcpyAsync(ib->vec[0], x, n4 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ib->fin, b, n4 * 8, hipMemcpyHostToDevice, c->stream));
  rc = ib_bicgstab_launch(c, ib->vec[0], ib->fin, tol, maxit, dt, rk_step, iters, err);
  if (rc && rc != C3D_ENOCONV) return rc;
  HIPCHK(c, hipMemcpyAsync(x, ib->vec[0], n4 * 8, hipMemcpyDeviceToHost, c->stream));
  const int s = sync(c);
  return s ? s : rc;
}

int c3d_ib_set_solver(c3d_ctx* c, double tol, int maxit, int guess_mode) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!(tol >= 0.0) || maxit < 0 || guess_mode < 0 || guess_mode > 5) return C3D_EINVAL;
  c->ib->tol = tol;
  c->ib->maxit = maxit;
  c->ib->guess = guess_mode;
  return C3D_OK;
}

int c3d_ib_set_graphs(c3d_ctx* c, int on) {
  int rc = need_ib(c);
  if (rc) return rc;
  c->ib->use_graphs = on != 0;
  return C3D_OK;
}

int c3d_ib_set_force(c3d_ctx* c, const double* fb) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!fb) return C3D_EINVAL;
  HIPCHK(c, hipMemcpyAsync(c->ib->fb, fb, 4 * (size_t)c->ib->m.nb * 8, hipMemcpyHostToDevice, c->stream));
  return sync(c);
}

int c3d_ib_get_force(c3d_ctx* c, double* fb) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!fb) return C3D_EINVAL;
  HIPCHK(c, hipMemcpyAsync(fb, c->ib->fb, 4 * (size_t)c->ib->m.nb * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_compute_ib_projection(c3d_ctx* c, double dt, int rk_step, int* iters, double* err) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!(dt != 0.0)) return fail(c, C3D_EINVAL, "dt must be nonzero");
  rc = ib_projection_launch(c, dt, rk_step);
  if (rc && rc != C3D_ENOCONV) return rc;
  if (iters) *iters = c->ib->last_iters;
  if (err) *err = c->ib->last_err;
  const int s = sync(c);
  return s ? s : rc;
}

int c3d_check_slip(c3d_ctx* c, double* max_slip) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!max_slip) return C3D_EINVAL;
  IBState* ib = c->ib;
  HIPCHK(c, hipMemsetAsync(c->red, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, c3d::ib::launch_max_slip(c->g, ib->m, c->U, c->V, c->W, ib->E1nH, ib->fb, c->red,
                                     c->stream));
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->red, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  if ((rc = sync(c))) return rc;
  std::memcpy(max_slip, &bits, sizeof(double));
  return C3D_OK;
}

int c3d_ib_stats(c3d_ctx* c, int* last_iters, long long* accum_iters, double* last_err) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (last_iters) *last_iters = c->ib->last_iters;
  if (accum_iters) *accum_iters = c->ib->accum_iters;
  if (last_err) *last_err = c->ib->last_err;
  return C3D_OK;
}

int c3d_ib_set_exterior_pressure_gradient(c3d_ctx* c, int flag) {
  int rc = need_ib(c);
  if (rc) return rc;
  IBState* ib = c->ib;
  ib->ext_pg = flag != 0;
  c->Hu = ib->ext_pg ? ib->Hu_int : nullptr;
  c->Hv = ib->ext_pg ? ib->Hv_int : nullptr;
  c->Hw = ib->ext_pg ? ib->Hw_int : nullptr;
  return C3D_OK;
}

// ----------------------------------------------------------------------------
// Learned initial guess for the Schur solve (no reference counterpart: the
// reference warm-starts BiCGSTAB from the previous force, src/projection.f90:60)
// ----------------------------------------------------------------------------
static uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Activation buffers for `rows` rows.  The larger set is allocated in full
// before the old one is released, so a failed allocation leaves the old,
// smaller set (and cap) intact and no pointer dangling.
static int mlp_reserve(c3d_ctx* c, int rows) {
  MLPState* m = c->mlp;
  if (rows <= m->cap) return 0;
  using namespace c3d::mlp;
  std::vector<void*> nb;
  auto get = [&](float** p, size_t n) -> bool {
    void* q = nullptr;
    if (hipMalloc(&q, n * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();  // the launchers report hipGetLastError()
      return false;
    }
    nb.push_back(q);
    *p = static_cast<float*>(q);
    return true;
  };
  float *X = nullptr, *T = nullptr, *dY = nullptr;
  float* act[kLayers] = {nullptr};
  float* dH[kLayers - 1] = {nullptr};
  bool ok = get(&X, (size_t)rows * kIn) && get(&T, (size_t)rows * kOut) &&
            get(&dY, (size_t)rows * kOut);
  for (int i = 0; i < kLayers && ok; i++) ok = get(&act[i], (size_t)rows * m->L.out[i]);
  for (int i = 0; i < kLayers - 1 && ok; i++) ok = get(&dH[i], (size_t)rows * m->L.out[i]);
  if (!ok) {
    for (void* p : nb) (void)hipFree(p);
    return fail(c, C3D_ENOMEM, "network buffers for %d rows", rows);
  }
  // no kernel may still read the old set (all work of the context is on its stream)
  (void)hipStreamSynchronize(c->stream);
  for (void* p : m->bufs) (void)hipFree(p);
  m->bufs.swap(nb);
  m->X = X;
  m->T = T;
  m->dY = dY;
  for (int i = 0; i < kLayers; i++) m->act[i] = act[i];
  for (int i = 0; i < kLayers - 1; i++) m->dH[i] = dH[i];
  m->cap = rows;
  return 0;
}

static int need_mlp(c3d_ctx* c) {
  if (!c) return C3D_EINVAL;
  if (!c->mlp) return fail(c, C3D_EINVAL, "no network (c3d_mlp_init)");
  return 0;
}

// Marker neighbourhoods on the slab layout: marker i of slab j of a ring of
// n markers has neighbours (i + di mod n, j + dj mod nzb), di, dj in -1..1.
static int mlp_neighbours(c3d_ctx* c) {
  MLPState* m = c->mlp;
  IBState* ib = c->ib;
  if (m->nb9) return 0;
  const int nb = ib->m.nb, nxb = ib->m.nxb, nxb1 = ib->m.nxb1, nzb = ib->m.nzb;
  std::vector<int> h((size_t)9 * nb);
  for (int j = 0; j < nzb; j++)
    for (int i = 0; i < nxb; i++) {
      const int lo = i < nxb1 ? 0 : nxb1, n = i < nxb1 ? nxb1 : nxb - nxb1, ii = i - lo;
      int s = 0;
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          const int jj = (j + dj + nzb) % nzb, i2 = lo + (ii + di + n) % n;
          h[(size_t)9 * (j * nxb + i) + s++] = jj * nxb + i2;
        }
    }
  int* d = nullptr;
  int rc = dalloc(c, &d, h.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  m->nb9 = d;
  return 0;
}

// Learned start of the Schur solve.  The network proposes an increment d of
// the force; the start is fb = fprev + a d with the step a minimising
// |b - S(fprev + a d)| (one Schur product S d beyond the warm start's
// S fprev), so its residual is never larger than the warm start's.  Leaves
// the residual b - S fb in vec[1] for BiCGSTAB.
static int mlp_guess(c3d_ctx* c, double dt, int rk_step) {
  MLPState* m = c->mlp;
  IBState* ib = c->ib;
  if (!m || !m->ready) return fail(c, C3D_EINVAL, "learned guess requested without a trained network");
  const int nb = ib->m.nb;
  const size_t n4 = 4 * (size_t)nb;
  int rc = mlp_neighbours(c);
  if (!rc) rc = mlp_reserve(c, nb);
  if (rc) return rc;
  namespace N = c3d::mlp;
  namespace B = c3d::ib;
  double *r = ib->vec[1], *d = ib->vec[5], *Sd = ib->vec[6];
  if ((rc = ib_schur_launch(c, ib->fb, r, dt, rk_step))) return rc;  // S fprev
  HIPCHK(c, B::launch_bicg_init(ib->rhs_ib, r, ib->vec[2], ib->vec[3], ib->vec[4], n4, c->stream));
  // per-solve normalisation: the residual channels by their rms over the 4nb
  // rows, the force channels by theirs over dudn_x and dudn_y (2nb)
  HIPCHK(c, B::launch_dot(r, r, nullptr, nullptr, n4, ib->part, 2, c->stream));
  HIPCHK(c, B::launch_dot(ib->fprev, ib->fprev, nullptr, nullptr, 2 * (size_t)nb, ib->part, 3,
                          c->stream));
  HIPCHK(c, B::launch_norm_scales(ib->part, 2, n4, 3, 2 * (size_t)nb, m->nrm, c->stream));
  HIPCHK(c, N::features(nb, m->nb9, r, ib->fprev, m->nrm, m->iscale, m->X, c->stream));
  HIPCHK(c, N::forward(m->L, m->P, m->X, nb, m->act, c->stream));
  HIPCHK(c, N::apply(nb, m->act[N::kLayers - 1], m->oscale, nullptr, d, c->stream));
  if ((rc = ib_schur_launch(c, d, Sd, dt, rk_step))) return rc;      // S d
  HIPCHK(c, B::launch_dot(r, Sd, Sd, Sd, n4, ib->part, 2, c->stream));
  HIPCHK(c, B::launch_bicg_scalar(5, ib->sc, ib->part, nullptr, c->stream));
  HIPCHK(c, B::launch_line_step(ib->fb, d, r, Sd, ib->sc, n4, c->stream));
  return 0;
}

int c3d_mlp_param_count(void) { return c3d::mlp::layout().total; }

int c3d_mlp_init(c3d_ctx* c, unsigned long long seed) {
  if (!c) return C3D_EINVAL;
  if (!c->mlp) {
    MLPState* m = new MLPState();
    m->L = c3d::mlp::layout();
    const size_t n = m->L.total;
    int rc = 0;
    c->mlp = m;
    if ((rc = dalloc(c, &m->P, n)) || (rc = dalloc(c, &m->G, n)) || (rc = dalloc(c, &m->M, n)) ||
        (rc = dalloc(c, &m->V, n)) || (rc = dalloc(c, &m->iscale, c3d::mlp::kCh)) ||
        (rc = dalloc(c, &m->oscale, c3d::mlp::kOut)) || (rc = dalloc(c, &m->ones, c3d::mlp::kCh)) ||
        (rc = dalloc(c, &m->part, c3d::mlp::kLossBlocks)) || (rc = dalloc(c, &m->nrm, 4)))
      return rc;
  }
  MLPState* m = c->mlp;
  // He-uniform weights (ReLU layers), zero biases
  std::vector<float> P(m->L.total, 0.f);
  uint64_t st = seed;
  for (int i = 0; i < c3d::mlp::kLayers; i++) {
    const float a = std::sqrt(6.f / (float)m->L.in[i]);
    for (int e = 0; e < m->L.out[i] * m->L.in[i]; e++) {
      const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      P[m->L.w[i] + e] = (float)((2.0 * u - 1.0) * a);
    }
  }
  const std::vector<float> one(8, 1.f);
  HIPCHK(c, hipMemcpy(m->P, P.data(), P.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(m->M, 0, P.size() * 4));
  HIPCHK(c, hipMemset(m->V, 0, P.size() * 4));
  HIPCHK(c, hipMemcpy(m->iscale, one.data(), c3d::mlp::kCh * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(m->oscale, one.data(), c3d::mlp::kOut * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(m->ones, one.data(), c3d::mlp::kCh * 4, hipMemcpyHostToDevice));
  const double unit[4] = {1.0, 1.0, 1.0, 1.0};  // nrm[2..3]: no normalisation (training pairs)
  HIPCHK(c, hipMemcpy(m->nrm, unit, sizeof(unit), hipMemcpyHostToDevice));
  m->t = 0;
  m->ready = false;
  return C3D_OK;
}

int c3d_mlp_set_scales(c3d_ctx* c, const float* iscale, const float* oscale) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!iscale || !oscale) return C3D_EINVAL;
  HIPCHK(c, hipMemcpy(c->mlp->iscale, iscale, c3d::mlp::kCh * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->mlp->oscale, oscale, c3d::mlp::kOut * 4, hipMemcpyHostToDevice));
  return C3D_OK;
}

int c3d_mlp_get_params(c3d_ctx* c, float* P) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!P) return C3D_EINVAL;
  HIPCHK(c, hipMemcpy(P, c->mlp->P, (size_t)c->mlp->L.total * 4, hipMemcpyDeviceToHost));
  return C3D_OK;
}

int c3d_mlp_set_params(c3d_ctx* c, const float* P) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!P) return C3D_EINVAL;
  HIPCHK(c, hipMemcpy(c->mlp->P, P, (size_t)c->mlp->L.total * 4, hipMemcpyHostToDevice));
  c->mlp->ready = true;
  return C3D_OK;
}

int c3d_mlp_forward(c3d_ctx* c, const float* X, int n, float* Y) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!X || !Y || n <= 0) return C3D_EINVAL;
  MLPState* m = c->mlp;
  namespace N = c3d::mlp;
  if ((rc = mlp_reserve(c, n))) return rc;
  HIPCHK(c, hipMemcpyAsync(m->X, X, (size_t)n * N::kIn * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, N::forward(m->L, m->P, m->X, n, m->act, c->stream));
  HIPCHK(c, hipMemcpyAsync(Y, m->act[N::kLayers - 1], (size_t)n * N::kOut * 4, hipMemcpyDeviceToHost,
                           c->stream));
  return sync(c);
}

static int mlp_loss_grad_launch(c3d_ctx* c, int n, float* loss) {
  MLPState* m = c->mlp;
  namespace N = c3d::mlp;
  HIPCHK(c, N::forward(m->L, m->P, m->X, n, m->act, c->stream));
  HIPCHK(c, N::loss_grad(m->act[N::kLayers - 1], m->T, n, m->dY, m->part, c->stream));
  HIPCHK(c, N::backward(m->L, m->P, m->X, n, m->act, m->dY, m->dH, m->G, c->stream));
  if (loss) {
    float h[N::kLossBlocks];
    HIPCHK(c, hipMemcpyAsync(h, m->part, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    int rc = sync(c);
    if (rc) return rc;
    double s = 0.0;
    for (int b = 0; b < N::kLossBlocks; b++) s += h[b];
    *loss = (float)s;
  }
  return 0;
}

int c3d_mlp_gradients(c3d_ctx* c, const float* X, const float* T, int n, float* G, float* loss) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!X || !T || !G || n <= 0) return C3D_EINVAL;
  MLPState* m = c->mlp;
  namespace N = c3d::mlp;
  if ((rc = mlp_reserve(c, n))) return rc;
  HIPCHK(c, hipMemcpyAsync(m->X, X, (size_t)n * N::kIn * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(m->T, T, (size_t)n * N::kOut * 4, hipMemcpyHostToDevice, c->stream));
  if ((rc = mlp_loss_grad_launch(c, n, loss))) return rc;
  HIPCHK(c, hipMemcpyAsync(G, m->G, (size_t)m->L.total * 4, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int c3d_mlp_train(c3d_ctx* c, const float* X, const float* T, int n, int epochs, int batch,
                  float lr, unsigned long long seed, float* loss_hist) {
  int rc = need_mlp(c);
  if (rc) return rc;
  if (!X || !T || n <= 0 || epochs <= 0 || batch <= 0 || batch % 16 || !(lr > 0.f))
    return fail(c, C3D_EINVAL, "training arguments (batch must be a positive multiple of 16)");
  MLPState* m = c->mlp;
  namespace N = c3d::mlp;
  const int nbatch = n / batch;
  if (nbatch == 0) return fail(c, C3D_EINVAL, "fewer samples than one batch");
  if ((rc = mlp_reserve(c, n > batch ? n : batch))) return rc;
  // shuffled copies of the data set, uploaded per epoch into the staging rows
  float *DX = nullptr, *DT = nullptr;
  if (hipMalloc(&DX, (size_t)n * N::kIn * 4) != hipSuccess ||
      hipMalloc(&DT, (size_t)n * N::kOut * 4) != hipSuccess) {
    (void)hipGetLastError();  // the launchers report hipGetLastError()
    if (DX) (void)hipFree(DX);
    return fail(c, C3D_ENOMEM, "training set buffers");
  }
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::vector<float> hx((size_t)n * N::kIn), ht((size_t)n * N::kOut);
  uint64_t st = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs && !rc; e++) {
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)(splitmix(st) % (uint64_t)(i + 1))]);
    for (int i = 0; i < n; i++) {
      std::memcpy(&hx[(size_t)i * N::kIn], &X[(size_t)perm[i] * N::kIn], N::kIn * 4);
      std::memcpy(&ht[(size_t)i * N::kOut], &T[(size_t)perm[i] * N::kOut], N::kOut * 4);
    }
    if (hipMemcpyAsync(DX, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(DT, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      rc = fail(c, C3D_EHIP, "training upload");
      break;
    }
    for (int b = 0; b < nbatch && !rc; b++) {
      if (hipMemcpyAsync(m->X, DX + (size_t)b * batch * N::kIn, (size_t)batch * N::kIn * 4,
                         hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
          hipMemcpyAsync(m->T, DT + (size_t)b * batch * N::kOut, (size_t)batch * N::kOut * 4,
                         hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        rc = fail(c, C3D_EHIP, "batch copy");
        break;
      }
      rc = mlp_loss_grad_launch(c, batch, nullptr);
      if (rc) break;
      m->t++;
      const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)m->t)) / (1.f - std::pow(b1, (float)m->t));
      if (N::adam(m->P, m->G, m->M, m->V, m->L.total, lr_t, b1, b2, eps, c->stream) != hipSuccess)
        rc = fail(c, C3D_EHIP, "adam");
    }
    if (!rc && loss_hist) {  // full-set loss after the epoch
      if (hipMemcpyAsync(m->X, DX, (size_t)n * N::kIn * 4, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
          hipMemcpyAsync(m->T, DT, (size_t)n * N::kOut * 4, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
          N::forward(m->L, m->P, m->X, n, m->act, c->stream) != hipSuccess ||
          N::loss_grad(m->act[N::kLayers - 1], m->T, n, m->dY, m->part, c->stream) != hipSuccess) {
        rc = fail(c, C3D_EHIP, "epoch loss");
        break;
      }
      float h[N::kLossBlocks];
      if (hipMemcpyAsync(h, m->part, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = fail(c, C3D_EHIP, "epoch loss read-back");
        break;
      }
      double s = 0.0;
      for (int q = 0; q < N::kLossBlocks; q++) s += h[q];
      loss_hist[e] = (float)s;
    }
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(DX);
  (void)hipFree(DT);
  if (!rc) m->ready = true;
  return rc;
}

int c3d_mlp_gemm_test(c3d_ctx* c, int M, int N, int K, int ta, int tb, const float* A, const float* B,
                      float* C) {
  if (!c || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return C3D_EINVAL;
  float *a = nullptr, *b = nullptr, *d = nullptr;
  int rc = 0;
  const size_t na = (size_t)M * K, nbv = (size_t)K * N, nc = (size_t)M * N;
  if (hipMalloc(&a, na * 4) != hipSuccess || hipMalloc(&b, nbv * 4) != hipSuccess ||
      hipMalloc(&d, nc * 4) != hipSuccess) {
    (void)hipGetLastError();  // the launchers report hipGetLastError()
    rc = fail(c, C3D_ENOMEM, "gemm test buffers");
  } else {
    const int lda = ta ? M : K, ldb = tb ? K : N;
    if (hipMemcpy(a, A, na * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b, B, nbv * 4, hipMemcpyHostToDevice) != hipSuccess ||
        c3d::mlp::gemm_test(M, N, K, ta, tb, a, lda, b, ldb, d, N, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(C, d, nc * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(c, C3D_EHIP, "gemm test");
  }
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (d) (void)hipFree(d);
  return rc;
}

int c3d_ib_last_pair(c3d_ctx* c, float* X, double* dF) {
  int rc = need_ib(c);
  if (rc) return rc;
  if (!X && !dF) return C3D_EINVAL;
  IBState* ib = c->ib;
  const int nb = ib->m.nb;
  namespace N = c3d::mlp;
  if (X) {
    if (!c->mlp && (rc = c3d_mlp_init(c, 1))) return rc;
    if ((rc = mlp_neighbours(c)) || (rc = mlp_reserve(c, nb))) return rc;
    // residual of the solve's warm start, b - S fprev (rk_step 2: the
    // operator without the pressure side effect)
    double* r = ib->vec[1];
    if ((rc = ib_schur_launch(c, ib->fprev, r, ib->last_dt, 2))) return rc;
    HIPCHK(c, c3d::ib::launch_bicg_init(ib->rhs_ib, r, ib->vec[2], ib->vec[3], ib->vec[4],
                                        4 * (size_t)nb, c->stream));
    HIPCHK(c, N::features(nb, c->mlp->nb9, r, ib->fprev, c->mlp->nrm + 2, c->mlp->ones, c->mlp->X,
                          c->stream));
    HIPCHK(c, hipMemcpyAsync(X, c->mlp->X, (size_t)nb * N::kIn * 4, hipMemcpyDeviceToHost, c->stream));
  }
  if (dF) {
    std::vector<double> a(4 * (size_t)nb), b(4 * (size_t)nb);
    HIPCHK(c, hipMemcpyAsync(a.data(), ib->fb, a.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.data(), ib->fprev, b.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if ((rc = sync(c))) return rc;
    for (size_t q = 0; q < a.size(); q++) dF[q] = a[q] - b[q];
  }
  return sync(c);
}

}  // extern "C"

// ---- Snapshots: the reference's big-endian stream file -----------------
// src/input_output.f90:498-728 (output_data) and :335-489 (read_input_data),
// one domain: every field is written whole, in the reference's global shape.

namespace {

struct BigEndianFile {
  FILE* f = nullptr;
  bool ok = true;
  ~BigEndianFile() {
    if (f) fclose(f);
  }
  static uint32_t swap32(uint32_t v) { return __builtin_bswap32(v); }
  static uint64_t swap64(uint64_t v) { return __builtin_bswap64(v); }
  void put_i4(int32_t v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u = swap32(u);
    ok = ok && fwrite(&u, 4, 1, f) == 1;
  }
  void put_f8s(const double* a, size_t n) {
    std::vector<uint64_t> buf(std::min<size_t>(n, 1 << 20));
    for (size_t o = 0; o < n && ok; o += buf.size()) {
      const size_t m = std::min(buf.size(), n - o);
      std::memcpy(buf.data(), a + o, m * 8);
      for (size_t i = 0; i < m; i++) buf[i] = swap64(buf[i]);
      ok = fwrite(buf.data(), 8, m, f) == m;
    }
  }
  void put_vec(const double* a, size_t n) {  // Write(1) Shape(a), a
    put_i4((int32_t)n);
    put_f8s(a, n);
  }
  bool get_i4(int32_t& v) {
    uint32_t u;
    if (fread(&u, 4, 1, f) != 1) return ok = false;
    u = swap32(u);
    std::memcpy(&v, &u, 4);
    return true;
  }
  bool get_f8s(double* a, size_t n) {
    if (fread(a, 8, n, f) != n) return ok = false;
    uint64_t* u = reinterpret_cast<uint64_t*>(a);
    for (size_t i = 0; i < n; i++) u[i] = swap64(u[i]);
    return true;
  }
  bool at_end() {
    const int ch = fgetc(f);
    if (ch == EOF) return true;
    ungetc(ch, f);
    return false;
  }
};

int download(c3d_ctx* c, const double* dev, size_t n, std::vector<double>& h) {
  h.resize(n);
  HIPCHK(c, hipMemcpyAsync(h.data(), dev, n * 8, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

// interior centres xm(i) = xg(i+1), i = 1..n-1 (src/initialization.f90:172-180)
std::vector<double> centres_of(const std::vector<double>& g) {
  return std::vector<double>(g.begin() + 1, g.end() - 1);
}

}  // namespace

extern "C" {

int c3d_write_snapshot(c3d_ctx* c, const char* path, double dt, const double* tangents_1,
                       const double* tangents_2) {
  if (!c || !path) return C3D_EINVAL;
  BigEndianFile out;
  out.f = fopen(path, "wb");
  if (!out.f) return fail(c, C3D_EINVAL, "cannot open %s for writing", path);
  // mesh (:520-527)
  out.put_vec(c->x.data(), c->x.size());
  out.put_vec(c->y.data(), c->y.size());
  out.put_vec(c->z.data(), c->z.size());
  const std::vector<double> xm = centres_of(c->xg), ym = centres_of(c->yg), zm = centres_of(c->zg);
  out.put_vec(xm.data(), xm.size());
  out.put_vec(ym.data(), ym.size());
  out.put_vec(zm.data(), zm.size());
  // U, V, W, P, then the interior Heaviside masks (:530-698)
  const int dims[7][3] = {{c->nx, c->nyg, c->nzg},  {c->nxg, c->ny, c->nzg}, {c->nxg, c->nyg, c->nz},
                          {c->nxg, c->nyg, c->nzg}, {c->nx, c->nyg, c->nzg}, {c->nxg, c->ny, c->nzg},
                          {c->nxg, c->nyg, c->nz}};
  IBState* ib = c->ib;
  const double* src[7] = {c->U, c->V, c->W, c->P, ib ? ib->Hu_int : nullptr,
                          ib ? ib->Hv_int : nullptr, ib ? ib->Hw_int : nullptr};
  std::vector<double> h;
  for (int a = 0; a < 7 && out.ok; a++) {
    const size_t n = (size_t)dims[a][0] * dims[a][1] * dims[a][2];
    int rc = 0;
    if (src[a]) {
      if ((rc = download(c, src[a], n, h))) return rc;
    } else {
      h.assign(n, 0.0);  // no body: the reference's masks are allocated, never set
    }
    for (int d = 0; d < 3; d++) out.put_i4(dims[a][d]);
    out.put_f8s(h.data(), n);
  }
  // t, dt, dpdx, nu and the body (:700-724)
  const double scal[4] = {c->t, dt, c->dPdx, c->desc.nu};
  out.put_f8s(scal, 4);
  const size_t nb = ib ? (size_t)ib->m.nb : 0;
  std::vector<double> mk[6], fb(4 * nb, 0.0);
  if (ib) {
    const double* msrc[6] = {ib->xb, ib->yb, ib->zb, ib->normals, ib->sb, ib->ub};
    const size_t mn[6] = {nb, nb, nb, 3 * nb, nb, 3 * nb};
    for (int a = 0; a < 6; a++) {
      int rc = download(c, msrc[a], mn[a], mk[a]);
      if (rc) return rc;
    }
    int rc = download(c, ib->fb, 4 * nb, fb);
    if (rc) return rc;
  }
  for (int a = 0; a < 3; a++) out.put_vec(mk[a].data(), mk[a].size());
  // nxb, nzb of the namelist: markers per slab (per wall for the wave bodies)
  const int nxb = !ib ? 0 : (ib->moving ? ib->wave.nxb : ib->m.nxb);
  out.put_i4(nxb);
  out.put_i4(ib ? ib->m.nzb : 0);
  out.put_vec(fb.data(), fb.size());
  out.put_vec(mk[5].data(), mk[5].size());  // ub
  out.put_vec(mk[4].data(), mk[4].size());  // sb
  out.put_vec(mk[3].data(), mk[3].size());  // normals
  const std::vector<double> zeros(3 * nb, 0.0);
  out.put_vec(tangents_1 ? tangents_1 : zeros.data(), 3 * nb);
  out.put_vec(tangents_2 ? tangents_2 : zeros.data(), 3 * nb);
  if (!out.ok) return fail(c, C3D_EINVAL, "write to %s failed", path);
  if (fclose(out.f) != 0) {
    out.f = nullptr;
    return fail(c, C3D_EINVAL, "closing %s failed", path);
  }
  out.f = nullptr;
  return C3D_OK;
}

int c3d_read_snapshot(c3d_ctx* c, const char* path, int what, double* dt_out) {
  if (!c || !path) return C3D_EINVAL;
  if (what & ~(C3D_RESTORE_P | C3D_RESTORE_STATE | C3D_RESTORE_FORCE))
    return fail(c, C3D_EINVAL, "unknown restore flags 0x%x", what);
  BigEndianFile in;
  in.f = fopen(path, "rb");
  if (!in.f) return fail(c, C3D_EINVAL, "cannot open %s", path);
  // mesh: sizes must match (:348-370); the coordinates must be this
  // context's grid (the library derives every operator from it)
  const std::vector<double> xm = centres_of(c->xg), ym = centres_of(c->yg), zm = centres_of(c->zg);
  const std::vector<double>* mesh[6] = {&c->x, &c->y, &c->z, &xm, &ym, &zm};
  const char* names[6] = {"x", "y", "z", "xm", "ym", "zm"};
  std::vector<double> h;
  for (int a = 0; a < 6; a++) {
    int32_t n = 0;
    if (!in.get_i4(n)) return fail(c, C3D_EINVAL, "%s: truncated mesh", path);
    if (n != (int32_t)mesh[a]->size())
      return fail(c, C3D_EINVAL, "%s: %s has %d points, the context %zu", path, names[a], n,
                  mesh[a]->size());
    h.resize(n);
    if (!in.get_f8s(h.data(), n)) return fail(c, C3D_EINVAL, "%s: truncated mesh", path);
    const double span = std::fabs(mesh[a]->back() - mesh[a]->front());
    for (int i = 0; i < n; i++)
      if (!(std::fabs(h[i] - (*mesh[a])[i]) <= 1e-12 * span))
        return fail(c, C3D_EINVAL, "%s: %s(%d) = %.17g differs from the context's grid (%.17g)",
                    path, names[a], i + 1, h[i], (*mesh[a])[i]);
  }
  const int dims[4][3] = {{c->nx, c->nyg, c->nzg}, {c->nxg, c->ny, c->nzg}, {c->nxg, c->nyg, c->nz},
                          {c->nxg, c->nyg, c->nzg}};
  std::vector<double> fld[4];
  for (int a = 0; a < 4; a++) {
    int32_t d[3];
    if (!in.get_i4(d[0]) || !in.get_i4(d[1]) || !in.get_i4(d[2]))
      return fail(c, C3D_EINVAL, "%s: truncated field header", path);
    if (d[0] != dims[a][0] || d[1] != dims[a][1] || d[2] != dims[a][2])
      return fail(c, C3D_EINVAL, "%s: wrong size in input file (%c)", path, "UVWP"[a]);
    fld[a].resize((size_t)d[0] * d[1] * d[2]);
    if (!in.get_f8s(fld[a].data(), fld[a].size()))
      return fail(c, C3D_EINVAL, "%s: truncated field", path);
  }
  double dt = 0.0, t = 0.0, dpdx = 0.0;
  std::vector<double> fb;
  const bool more = (what & (C3D_RESTORE_STATE | C3D_RESTORE_FORCE)) || dt_out;
  if (more) {
    if (in.at_end()) {
      if (what & (C3D_RESTORE_STATE | C3D_RESTORE_FORCE))
        return fail(c, C3D_EINVAL, "%s: no time/body section (fields-only format)", path);
    } else {
      for (int a = 0; a < 3; a++) {  // Hu, Hv, Hw interior: recomputed from the body
        int32_t d[3];
        if (!in.get_i4(d[0]) || !in.get_i4(d[1]) || !in.get_i4(d[2]))
          return fail(c, C3D_EINVAL, "%s: truncated mask header", path);
        if (d[0] < 0 || d[1] < 0 || d[2] < 0 ||
            fseek(in.f, (long)((size_t)d[0] * d[1] * d[2] * 8), SEEK_CUR) != 0)
          return fail(c, C3D_EINVAL, "%s: truncated or corrupt mask", path);
      }
      double sc[4];
      if (!in.get_f8s(sc, 4)) return fail(c, C3D_EINVAL, "%s: truncated scalars", path);
      t = sc[0];
      dt = sc[1];
      dpdx = sc[2];
      // marker counts come from the file: bounded before any seek or allocation
      const int32_t kMaxLen = 1 << 28;
      int32_t n = 0;
      for (int a = 0; a < 3; a++) {  // xb, yb, zb
        if (!in.get_i4(n) || n < 0 || n > kMaxLen || fseek(in.f, (long)n * 8, SEEK_CUR) != 0)
          return fail(c, C3D_EINVAL, "%s: truncated or corrupt markers", path);
      }
      int32_t nxb = 0, nzb = 0;
      if (!in.get_i4(nxb) || !in.get_i4(nzb) || !in.get_i4(n) || n < 0 || n > kMaxLen)
        return fail(c, C3D_EINVAL, "%s: truncated or corrupt body section", path);
      fb.resize(n);
      if (!in.get_f8s(fb.data(), n)) return fail(c, C3D_EINVAL, "%s: truncated fb", path);
    }
  }
  bool force = (what & C3D_RESTORE_FORCE) != 0;
  if (force) {  // a body-free file into a body-free context has no force to restore
    if (!c->ib && !fb.empty())
      return fail(c, C3D_EINVAL, "%s holds a body force: set the body up before reading", path);
    if (c->ib && fb.size() != 4 * (size_t)c->ib->m.nb)
      return fail(c, C3D_EINVAL, "%s: fb has %zu entries, the body 4*%d", path, fb.size(),
                  c->ib->m.nb);
    force = c->ib != nullptr;
  }
  // U, V, W always (the reference's restart); the rest on request
  double* dst[4] = {c->U, c->V, c->W, c->P};
  for (int a = 0; a < 4; a++) {
    if (a == 3 && !(what & C3D_RESTORE_P)) break;
    HIPCHK(c, hipMemcpyAsync(dst[a], fld[a].data(), fld[a].size() * 8, hipMemcpyHostToDevice,
                             c->stream));
  }
  if (force)
    HIPCHK(c, hipMemcpyAsync(c->ib->fb, fb.data(), fb.size() * 8, hipMemcpyHostToDevice, c->stream));
  int rc = sync(c);
  if (rc) return rc;
  if (what & C3D_RESTORE_STATE) {
    c->t = t;
    c->dPdx = dpdx;
    if (c->ib && c->ib->moving && (rc = ib_move_to(c, t))) return rc;
  }
  if (dt_out) *dt_out = dt;
  return C3D_OK;
}

}  // extern "C"
