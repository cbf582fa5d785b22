// Host orchestration of the RS-AMG hierarchy: setup, numeric update and V-cycle.
#include "amg.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "amg_kernels.h"
#include "ktimer.h"
#include "prims.h"

namespace impes {

namespace {

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
void d2d(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) IMPES_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

// ---------------------------------------------------------------- stencil kernels
__device__ __forceinline__ void cell_ijk(long long c, int ny, int nz, int& i, int& j, int& k) {
  k = (int)(c % nz);
  const long long t = c / nz;
  j = (int)(t % ny);
  i = (int)(t / ny);
}

__global__ void k_st7_count(int nx, int ny, int nz, int* __restrict__ cnt) {
  const long long n = (long long)nx * ny * nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, ny, nz, i, j, k);
    cnt[c] = 1 + (i > 0) + (j > 0) + (k > 0) + (k < nz - 1) + (j < ny - 1) + (i < nx - 1);
  }
}

__global__ void k_st7_fill(int nx, int ny, int nz, const int* __restrict__ rp, int* __restrict__ ci) {
  const long long n = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz, sy = nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, ny, nz, i, j, k);
    int e = rp[c];
    if (i > 0) ci[e++] = (int)(c - sx);
    if (j > 0) ci[e++] = (int)(c - sy);
    if (k > 0) ci[e++] = (int)(c - 1);
    ci[e++] = (int)c;
    if (k < nz - 1) ci[e++] = (int)(c + 1);
    if (j < ny - 1) ci[e++] = (int)(c + sy);
    if (i < nx - 1) ci[e++] = (int)(c + sx);
  }
}

__global__ void k_st7_vals(int nx, int ny, int nz, const double* __restrict__ d, const double* __restrict__ ox,
                           const double* __restrict__ oy, const double* __restrict__ oz, const int* __restrict__ rp,
                           double* __restrict__ v) {
  const long long n = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz, sy = nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, ny, nz, i, j, k);
    int e = rp[c];
    if (i > 0) v[e++] = ox[c - sx];
    if (j > 0) v[e++] = oy[c - sy];
    if (k > 0) v[e++] = oz[c - 1];
    v[e++] = d[c];
    if (k < nz - 1) v[e++] = oz[c];
    if (j < ny - 1) v[e++] = oy[c];
    if (i < nx - 1) v[e++] = ox[c];
  }
}

// equilibrated stencil (and its FP32 copy): a_ij * dh_i * dh_j
__global__ void k_st7_scale(int nx, int ny, int nz, const double* __restrict__ d, const double* __restrict__ ox,
                            const double* __restrict__ oy, const double* __restrict__ oz,
                            const double* __restrict__ dh, double* __restrict__ sd, double* __restrict__ sxo,
                            double* __restrict__ syo, double* __restrict__ szo, float* __restrict__ sd32,
                            float* __restrict__ sx32, float* __restrict__ sy32, float* __restrict__ sz32) {
  const long long n = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz, sy = nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, ny, nz, i, j, k);
    const double h = dh[c];
    const double vd = d[c] * h * h;
    const double vx = (i < nx - 1) ? ox[c] * h * dh[c + sx] : 0.0;
    const double vy = (j < ny - 1) ? oy[c] * h * dh[c + sy] : 0.0;
    const double vz = (k < nz - 1) ? oz[c] * h * dh[c + 1] : 0.0;
    sd[c] = vd; sxo[c] = vx; syo[c] = vy; szo[c] = vz;
    sd32[c] = (float)vd; sx32[c] = (float)vx; sy32[c] = (float)vy; sz32[c] = (float)vz;
  }
}

// Per-step refresh of level 0 from a new stencil, one pass: the equilibrated CSR
// values (a_ij * dh_i * dh_j, as k_sym_scale forms them), the L1 relaxation
// weights (the row sum folded in k_diag_relax's eight-lane order, so the weights
// equal that kernel's bit for bit), the equilibrated stencil and its FP32 copy
// (as k_st7_scale), the input copies, and the preconditioner's scalings. Each
// block stages the CSR values of its 256 rows in LDS and writes their contiguous
// range coalesced.
struct St7FineUpdate {
  int nx, ny, nz;
  const double *d, *ox, *oy, *oz;  // unscaled stencil, as given
  const double* dh;                // diagonal equilibration factors (null: every factor is a)
  double a;
  const int* rp;
  double *st_d, *st_x, *st_y, *st_z, *diag0;
  double* v;
  double* ir64;
  float* ir32;
  double *sd, *sx, *sy, *sz;
  float *sd32, *sx32, *sy32, *sz32;
  double *dh_out, *sin, *sout;     // dh_out written only when every factor is a
};

constexpr int kFineRows = 256;

template <bool kScalar>
__global__ __launch_bounds__(kFineRows) void k_st7_fine_update(St7FineUpdate u) {
  __shared__ double stage[kFineRows * 7];
  const long long n = (long long)u.nx * u.ny * u.nz;
  const long long sx = (long long)u.ny * u.nz, sy = u.nz;
  for (long long c0 = (long long)blockIdx.x * kFineRows; c0 < n; c0 += (long long)gridDim.x * kFineRows) {
    const long long c = c0 + threadIdx.x;
    const long long c1 = (c0 + kFineRows < n) ? c0 + kFineRows : n;
    const int base = u.rp[c0];
    if (c < n) {
      int i, j, k;
      cell_ijk(c, u.ny, u.nz, i, j, k);
      const double h = kScalar ? u.a : u.dh[c];
      const double dc = u.d[c], xc = u.ox[c], yc = u.oy[c], zc = u.oz[c];
      // CSR order of a row: -x, -y, -z, diagonal, +z, +y, +x
      const bool pr[7] = {i > 0, j > 0, k > 0, true, k < u.nz - 1, j < u.ny - 1, i < u.nx - 1};
      double e[7];
      e[0] = pr[0] ? u.ox[c - sx] * h * (kScalar ? u.a : u.dh[c - sx]) : 0.0;
      e[1] = pr[1] ? u.oy[c - sy] * h * (kScalar ? u.a : u.dh[c - sy]) : 0.0;
      e[2] = pr[2] ? u.oz[c - 1] * h * (kScalar ? u.a : u.dh[c - 1]) : 0.0;
      e[3] = dc * h * h;
      e[4] = pr[4] ? zc * h * (kScalar ? u.a : u.dh[c + 1]) : 0.0;
      e[5] = pr[5] ? yc * h * (kScalar ? u.a : u.dh[c + sy]) : 0.0;
      e[6] = pr[6] ? xc * h * (kScalar ? u.a : u.dh[c + sx]) : 0.0;
      // |a_ij| of the row's m-th entry sits in lane m of k_diag_relax's group
      double l[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int m = 0;
#pragma unroll
      for (int t = 0; t < 7; ++t) {
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (pr[t] && m == q) l[q] = fabs(e[t]);
        m += pr[t] ? 1 : 0;
      }
      const double rs = ((l[0] + l[4]) + (l[2] + l[6])) + ((l[1] + l[5]) + (l[3] + l[7]));
      double den = fmax(fabs(e[3]), 0.3 * rs);
      den = fmax(den, 1e-30);
      const double ir = fmin(1.0 / den, 1e2);
      u.ir64[c] = ir;
      u.ir32[c] = (float)ir;
      int p = u.rp[c] - base;
#pragma unroll
      for (int t = 0; t < 7; ++t)
        if (pr[t]) stage[p++] = e[t];
      u.st_d[c] = dc; u.st_x[c] = xc; u.st_y[c] = yc; u.st_z[c] = zc; u.diag0[c] = dc;
      u.sd[c] = e[3]; u.sx[c] = e[6]; u.sy[c] = e[5]; u.sz[c] = e[4];
      u.sd32[c] = (float)e[3]; u.sx32[c] = (float)e[6]; u.sy32[c] = (float)e[5]; u.sz32[c] = (float)e[4];
      if (kScalar) u.dh_out[c] = h;
      u.sin[c] = h * dc;
      u.sout[c] = h;
    }
    __syncthreads();
    const int cnt = u.rp[c1] - base;
    for (int q = threadIdx.x; q < cnt; q += kFineRows) u.v[base + q] = stage[q];
    __syncthreads();
  }
}

__global__ void k_set_cf(signed char* cf, int idx) { cf[idx] = kC; }
__global__ void k_neg_flags(int m, const int* __restrict__ tp, int* __restrict__ out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < m; e += (long long)gridDim.x * blockDim.x)
    out[e] = tp[e] < 0 ? 1 : 0;
}

// ---------------------------------------------------------------- precision selection
template <class T>
struct Sel;
template <>
struct Sel<double> {
  static double* x(Level& L) { return L.x64.p; }
  static double* b(Level& L) { return L.b64.p; }
  static double* r(Level& L) { return L.r64.p; }
  static double* t(Level& L) { return L.t64.p; }
  static const double* ir(const Level& L) { return L.ir64.p; }
  static CsrF64 A(const Level& L) { return CsrF64{L.n, L.n, L.nnz, L.rp.p, L.ci.p, L.v.p}; }
  static CsrF64 P(const Level& L) { return CsrF64{L.n, L.nc, L.pnnz, L.prp.p, L.pci.p, L.pv.p}; }
  static CsrF64 R(const Level& L) { return CsrF64{L.nc, L.n, L.pnnz, L.rrp.p, L.rci.p, L.rv.p}; }
  static void jac(const CsrF64& A, const double* xi, const double* b, const double* ir, double w, double* xo, int g,
                  hipStream_t s) { launch_jacobi_f64(A, xi, b, ir, w, xo, g, s); }
  static void jac0(const double* b, const double* ir, double w, double* xo, int n, hipStream_t s) {
    launch_jacobi0_f64(b, ir, w, xo, n, s);
  }
  static void jacL(const Level& L, const double* xi, const double* b, double w, double* xo, hipStream_t s) {
    launch_jacobi_f64(A(L), xi, b, L.ir64.p, w, xo, L.group, s);
  }
  static void resL(const Level& L, const double* x, const double* b, double* r, hipStream_t s) {
    launch_residual_f64(A(L), x, b, r, L.group, s);
  }
  static void res(const CsrF64& A, const double* x, const double* b, double* r, int g, hipStream_t s) {
    launch_residual_f64(A, x, b, r, g, s);
  }
  static void spmv(const CsrF64& A, const double* x, double* y, int g, hipStream_t s) { launch_spmv_f64(A, x, y, g, s); }
  static void spmv_add(const CsrF64& A, const double* x, double* y, int g, hipStream_t s) {
    launch_spmv_add_f64(A, x, y, g, s);
  }
  static void spmv_add_to(const CsrF64& A, const double* x, const double* y0, double* y, int g, hipStream_t s) {
    launch_spmv_add_to_f64(A, x, y0, y, g, s);
  }
  static void copy(double* y, const double* x, long long n, hipStream_t s) { launch_copy_f64(y, x, n, s); }
  static void rj0(const CsrF64& R, const double* r, const double* ir, double w, double* bc, double* xc, int g,
                  hipStream_t s) { launch_restrict_j0_f64(R, r, ir, w, bc, xc, g, s); }
  static void pin0(const int* pin, double* b, hipStream_t s) { launch_pin_rhs_f64(pin, b, s); }
  // grid transfers of level L (P: n x nc, R: nc x n)
  static void restrict_j0(const Level& L, const double* r, const double* irc, double w, double* bc, double* xc,
                          hipStream_t s) { launch_restrict_j0_f64(R(L), r, irc, w, bc, xc, L.rgroup, s); }
  static void restrict_to(const Level& L, const double* r, double* bc, hipStream_t s) {
    launch_spmv_f64(R(L), r, bc, L.rgroup, s);
  }
  static void prolong(const Level& L, const double* e, double* y, hipStream_t s) {
    launch_spmv_f64(P(L), e, y, L.pgroup, s);
  }
  static void prolong_add(const Level& L, const double* e, double* y, hipStream_t s) {
    launch_spmv_add_f64(P(L), e, y, L.pgroup, s);
  }
  static void prolong_add_to(const Level& L, const double* e, const double* y0, double* y, hipStream_t s) {
    launch_spmv_add_to_f64(P(L), e, y0, y, L.pgroup, s);
  }
  static void dense_mv(int n, const int* pin, const double* M, const double* b, double* x, hipStream_t s) {
    launch_dense_mv_f64(n, pin, M, b, x, s);
  }
  static void dense(int n, const int* pin, const double* aug, const double* b, double* x, hipStream_t s) {
    launch_dense_solve_f64(n, pin, aug, b, x, s);
  }
};
template <>
struct Sel<float> {
  static float* x(Level& L) { return L.x32.p; }
  static float* b(Level& L) { return L.b32.p; }
  static float* r(Level& L) { return L.r32.p; }
  static float* t(Level& L) { return L.t32.p; }
  static const float* ir(const Level& L) { return L.ir32.p; }
  static CsrF32 A(const Level& L) { return CsrF32{L.n, L.n, L.nnz, L.rp.p, L.ci.p, L.v32.p}; }
  static CsrF32 P(const Level& L) { return CsrF32{L.n, L.nc, L.pnnz, L.prp.p, L.pci.p, L.pv32.p}; }
  static CsrF32 R(const Level& L) { return CsrF32{L.nc, L.n, L.pnnz, L.rrp.p, L.rci.p, L.rv32.p}; }
  static void jac(const CsrF32& A, const float* xi, const float* b, const float* ir, double w, float* xo, int g,
                  hipStream_t s) { launch_jacobi_f32(A, xi, b, ir, (float)w, xo, g, s); }
  static void jac0(const float* b, const float* ir, double w, float* xo, int n, hipStream_t s) {
    launch_jacobi0_f32(b, ir, (float)w, xo, n, s);
  }
  static Csr4F32 A4(const Level& L) {
    return Csr4F32{L.n, L.nnz, L.rp4.p, reinterpret_cast<const int4*>(L.ci4.p), reinterpret_cast<const float4*>(L.v4.p),
                   reinterpret_cast<const short4*>(L.co4.p), L.nw4, L.wrow4.p, L.wq4.p};
  }
  static void jacL(const Level& L, const float* xi, const float* b, double w, float* xo, hipStream_t s) {
    if (L.pad4) launch_jacobi4_f32(A4(L), L.g4, xi, b, L.ir32.p, (float)w, xo, s);
    else launch_jacobi_f32(A(L), xi, b, L.ir32.p, (float)w, xo, L.group, s);
  }
  static void resL(const Level& L, const float* x, const float* b, float* r, hipStream_t s) {
    if (L.pad4) launch_residual4_f32(A4(L), L.g4, x, b, r, s);
    else launch_residual_f32(A(L), x, b, r, L.group, s);
  }
  static void res(const CsrF32& A, const float* x, const float* b, float* r, int g, hipStream_t s) {
    launch_residual_f32(A, x, b, r, g, s);
  }
  static void spmv(const CsrF32& A, const float* x, float* y, int g, hipStream_t s) { launch_spmv_f32(A, x, y, g, s); }
  static void spmv_add(const CsrF32& A, const float* x, float* y, int g, hipStream_t s) {
    launch_spmv_add_f32(A, x, y, g, s);
  }
  static void spmv_add_to(const CsrF32& A, const float* x, const float* y0, float* y, int g, hipStream_t s) {
    launch_spmv_add_to_f32(A, x, y0, y, g, s);
  }
  static void copy(float* y, const float* x, long long n, hipStream_t s) { launch_copy_f32(y, x, n, s); }
  static void rj0(const CsrF32& R, const float* r, const float* ir, double w, float* bc, float* xc, int g,
                  hipStream_t s) { launch_restrict_j0_f32(R, r, ir, (float)w, bc, xc, g, s); }
  static void pin0(const int* pin, float* b, hipStream_t s) { launch_pin_rhs_f32(pin, b, s); }
  // grid transfers of level L: restriction in the padded tiled form when built
  static Csr4F32 T4(const Pad4& T) {
    return Csr4F32{T.n, T.nnz, T.rp4.p, reinterpret_cast<const int4*>(T.ci4.p), reinterpret_cast<const float4*>(T.v4.p),
                   nullptr, T.nw, T.wrow.p, T.wq.p};
  }
  static void restrict_j0(const Level& L, const float* r, const float* irc, double w, float* bc, float* xc,
                          hipStream_t s) {
    if (L.r4.nw) launch_restrict4_j0_f32(T4(L.r4), r, irc, (float)w, bc, xc, s);
    else launch_restrict_j0_f32(R(L), r, irc, (float)w, bc, xc, L.rgroup, s);
  }
  static void restrict_to(const Level& L, const float* r, float* bc, hipStream_t s) {
    if (L.r4.nw) launch_spmv4_f32(T4(L.r4), r, bc, s);
    else launch_spmv_f32(R(L), r, bc, L.rgroup, s);
  }
  static void prolong(const Level& L, const float* e, float* y, hipStream_t s) {
    launch_spmv_f32(P(L), e, y, L.pgroup, s);
  }
  static void prolong_add(const Level& L, const float* e, float* y, hipStream_t s) {
    launch_spmv_add_f32(P(L), e, y, L.pgroup, s);
  }
  static void prolong_add_to(const Level& L, const float* e, const float* y0, float* y, hipStream_t s) {
    launch_spmv_add_to_f32(P(L), e, y0, y, L.pgroup, s);
  }
  static void dense_mv(int n, const int* pin, const double* M, const float* b, float* x, hipStream_t s) {
    launch_dense_mv_f32(n, pin, M, b, x, s);
  }
  static void dense(int n, const int* pin, const double* aug, const float* b, float* x, hipStream_t s) {
    launch_dense_solve_f32(n, pin, aug, b, x, s);
  }
};

}  // namespace

void launch_st7_pattern_count(int nx, int ny, int nz, int* cnt, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  k_st7_count<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(nx, ny, nz, cnt);
  IMPES_LAUNCH_CHECK();
}
void launch_st7_pattern_fill(int nx, int ny, int nz, const int* rp, int* ci, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  k_st7_fill<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(nx, ny, nz, rp, ci);
  IMPES_LAUNCH_CHECK();
}
void launch_st7_to_csr_values(int nx, int ny, int nz, const double* d, const double* ox, const double* oy,
                              const double* oz, const int* rp, double* v, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  k_st7_vals<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(nx, ny, nz, d, ox, oy, oz, rp, v);
  IMPES_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- setup
void Hierarchy::setup_csr(int n, int nnz, const int* rp, const int* ci, const double* v, hipStream_t s) {
  stencil_ = false;
  setup_impl(n, nnz, rp, ci, v, s);
}

Hierarchy::Hierarchy(const AmgParams& p) : prm_(p) {
  // IMPES_GRAPHS=0 runs every V-cycle as eager launches
  const char* g = std::getenv("IMPES_GRAPHS");
  graphs_ = !(g && g[0] == '0');
}

Hierarchy::~Hierarchy() { drop_graphs(); }

namespace {
// Capture-only stream (nothing runs on it): one per thread and device, created on
// first use and kept until the process ends. Creating a stream takes ~4 ms, which
// every new hierarchy (each environment reset, for one) would otherwise pay.
hipStream_t capture_stream() {
  thread_local std::map<int, hipStream_t> streams;
  int dev = 0;
  IMPES_HIP_CHECK(hipGetDevice(&dev));
  auto it = streams.find(dev);
  if (it != streams.end()) return it->second;
  hipStream_t cs = nullptr;
  IMPES_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  streams[dev] = cs;
  return cs;
}
}  // namespace

void Hierarchy::drop_graphs() {
  for (auto& e : vexec_) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

void Hierarchy::capture_vcycle(int precision) {
  hipStream_t cs = capture_stream();
  Level& L0 = levels_[0];
  hipGraph_t g = nullptr;
  IMPES_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  try {
    if (precision == 32) vcycle_level<float>(0, L0.b32.p, L0.x32.p, true, cs);
    else vcycle_level<double>(0, L0.b64.p, L0.x64.p, true, cs);
  } catch (...) {
    (void)hipStreamEndCapture(cs, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  IMPES_HIP_CHECK(hipStreamEndCapture(cs, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  IMPES_HIP_CHECK(rc);
  vexec_[precision == 32 ? 0 : 1] = e;
}

void Hierarchy::setup_impl(int n, int nnz, const int* rp, const int* ci, const double* v, hipStream_t s) {
  IMPES_REQUIRE(n >= 1 && nnz >= 1, "amg setup: empty matrix");
  const double t0 = wall();
  {
  // every launch and copy of the build goes to s; the scope closes after the
  // stream synchronisation below (on an error it frees the deferred blocks
  // with hipFree, which waits for the device)
  PoolScope scope(pool_);
  drop_graphs();  // the level buffers are about to be rebuilt
  levels_.clear();
  levels_.emplace_back();
  Level& L0 = levels_[0];
  L0.n = n;
  L0.nnz = nnz;
  L0.rp.alloc(n + 1);
  L0.ci.alloc(nnz);
  L0.v.alloc(nnz);
  d2d(L0.rp.p, rp, n + 1, s);
  d2d(L0.ci.p, ci, nnz, s);
  v0_unscaled_.alloc(nnz);
  d2d(v0_unscaled_.p, v, nnz, s);
  equilibrate(v0_unscaled_.p, true, s);
  build_levels(s);
  refresh_f32(s);
  if (stencil_) refresh_stencil(s);
  alloc_vectors();
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  }
  last_setup_seconds = wall() - t0;
}

void Hierarchy::setup_stencil(int nx, int ny, int nz, const double* d, const double* ox, const double* oy,
                              const double* oz, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  IMPES_REQUIRE(n >= 1 && n < (1LL << 30), "amg setup: bad stencil grid");
  stencil_ = true;
  nx_ = nx; ny_ = ny; nz_ = nz;
  st_d_.alloc(n); st_x_.alloc(n); st_y_.alloc(n); st_z_.alloc(n);
  d2d(st_d_.p, d, n, s); d2d(st_x_.p, ox, n, s); d2d(st_y_.p, oy, n, s); d2d(st_z_.p, oz, n, s);
  DevBuf<int> cnt, rp, ci;
  DevBuf<double> v;
  cnt.alloc(n);
  rp.alloc(n + 1);
  launch_st7_pattern_count(nx, ny, nz, cnt.p, s);
  scan_i32_to_i32(cnt.p, rp.p, n, s);
  const int nnz = read_scalar(rp.p + n, s);
  ci.alloc(nnz);
  v.alloc(nnz);
  launch_st7_pattern_fill(nx, ny, nz, rp.p, ci.p, s);
  launch_st7_to_csr_values(nx, ny, nz, d, ox, oy, oz, rp.p, v.p, s);
  setup_impl((int)n, nnz, rp.p, ci.p, v.p, s);
}

void Hierarchy::equil_factors(const double* diag, bool decide, hipStream_t s) {
  const int n = levels_[0].n;
  if (prm_.equil_mode == kEquilScalar) {
    if (decide) {
      // lower median of |a_ii| (torch.median semantics), chosen once per setup
      std::vector<double> d((size_t)n);
      IMPES_HIP_CHECK(hipMemcpyAsync(d.data(), diag, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      IMPES_HIP_CHECK(hipStreamSynchronize(s));
      for (auto& x : d) x = std::max(std::fabs(x), 1e-30);
      auto mid = d.begin() + (n - 1) / 2;
      std::nth_element(d.begin(), mid, d.end());
      equil_ = *mid < prm_.equil_threshold;
      scalar_ = equil_ ? 1.0 / std::sqrt(*mid) : 1.0;
    }
    launch_fill_f64(dh_.p, scalar_, n, s);
  } else {
    DevBuf<int> flags;
    flags.alloc(n);
    launch_count_small_diag(n, diag, prm_.equil_threshold, flags.p, s);
    // lower median (torch.median semantics) below the threshold  <=>  at least
    // (n-1)/2 + 1 entries below it
    const long long small = sum_i32(flags.p, n, s);
    equil_ = small >= (long long)((n - 1) / 2 + 1);
    if (equil_) launch_dhalf_inv(n, diag, dh_.p, s);
    else launch_fill_f64(dh_.p, 1.0, n, s);
    sync_before_release(s);  // flags freed on return
  }
}

void Hierarchy::equilibrate(const double* vu, bool decide, hipStream_t s) {
  Level& L0 = levels_[0];
  const int n = L0.n;
  diag0_.alloc(n);
  dh_.alloc(n);
  sin_.alloc(n);
  sout_.alloc(n);
  launch_diag_relax(n, L0.rp.p, L0.ci.p, vu, diag0_.p, nullptr, nullptr, s);
  equil_factors(diag0_.p, decide, s);
  // unequilibrated: every factor is 1 and the scaled values equal the given ones
  if (equil_) launch_sym_scale(n, L0.rp.p, L0.ci.p, vu, dh_.p, L0.v.p, s);
  else d2d(L0.v.p, vu, L0.nnz, s);
  launch_scale_f64(sin_.p, dh_.p, diag0_.p, n, s);
  launch_copy_f64(sout_.p, dh_.p, n, s);
}

void Hierarchy::build_levels(hipStream_t s) {
  // IMPES_AMG_TRACE=1 prints one line per level built (size, pattern, seconds) to stderr
  const char* tr = std::getenv("IMPES_AMG_TRACE");
  const bool trace = tr && tr[0] == '1';
  double tl = wall();
  for (int lvl = 0; lvl < prm_.max_levels; ++lvl) {
    const int n = levels_[lvl].n;
    if (n <= prm_.coarsest_size) {
      levels_[lvl].coarsest = true;
      break;
    }
    Level& A = levels_[lvl];
    const int nnz = A.nnz;
    DevBuf<unsigned char> strong;
    DevBuf<int> tpos, deg, is_c, f2c, cnt_c, cand, pcnt, oc, colcnt, cursor, rrp32;
    DevBuf<signed char> is_max, cf0;
    DevBuf<double> sum_strong, cvals, ov;
    DevBuf<long long> cand_ptr, rrp64;
    DevBuf<unsigned long long> keys;
    strong.alloc(nnz);
    tpos.alloc(nnz);
    launch_strength(n, A.nnz, A.rp.p, A.ci.p, A.v.p, prm_.theta, strong.p, s);
    launch_transpose_pos(n, A.nnz, A.rp.p, A.ci.p, tpos.p, s);
    {
      // structural symmetry is required by the symmetrised strength graph
      DevBuf<int> neg;
      neg.alloc(nnz);
      k_neg_flags<<<grid_for(nnz, 256, 1 << 16), 256, 0, s>>>(nnz, tpos.p, neg.p);
      IMPES_LAUNCH_CHECK();
      IMPES_REQUIRE(sum_i32(neg.p, nnz, s) == 0,
                    "amg setup: matrix pattern is not structurally symmetric");
    }
    deg.alloc(n);
    is_max.alloc(n);
    cf0.alloc(n);
    if (prm_.coarsening == kCoarsenIterated) {
      DevBuf<int> undecided;
      undecided.alloc(n);
      launch_cf_split_iterated(n, A.nnz, A.rp.p, A.ci.p, strong.p, tpos.p, deg.p, is_max.p, undecided.p, cf0.p, s);
    } else {
      launch_cf_split(n, A.nnz, A.rp.p, A.ci.p, strong.p, tpos.p, deg.p, is_max.p, cf0.p, s);
    }
    if (prm_.dominant > 0.0) launch_force_dominant(n, A.rp.p, A.ci.p, A.v.p, prm_.dominant, cf0.p, s);
    if (prm_.anchor >= 0 && prm_.anchor < n) {
      k_set_cf<<<1, 1, 0, s>>>(cf0.p, prm_.anchor);
      IMPES_LAUNCH_CHECK();
    }
    const long long n_coarse = count_i8(cf0.p, n, kC, s);
    if ((double)n_coarse >= 0.9 * n) {
      A.coarsest = true;
      break;
    }
    // orphan promotion and coarse numbering
    A.cf.alloc(n);
    is_c.alloc(n);
    launch_orphans(n, A.rp.p, A.ci.p, strong.p, cf0.p, A.cf.p, is_c.p, s);
    f2c.alloc(n + 1);
    scan_i32_to_i32(is_c.p, f2c.p, n, s);
    const int nc = read_scalar(f2c.p + n, s);
    IMPES_REQUIRE(nc >= 1, "amg setup: empty coarse level");
    // interpolation
    sum_strong.alloc(n);
    cnt_c.alloc(n);
    cand.alloc(n);
    launch_interp_pass_a(n, A.rp.p, A.ci.p, A.v.p, strong.p, A.cf.p, sum_strong.p, cnt_c.p, s);
    launch_interp_count(n, A.rp.p, A.ci.p, strong.p, A.cf.p, sum_strong.p, cnt_c.p, cand.p, s);
    cand_ptr.alloc(n + 1);
    scan_i32_to_i64(cand.p, cand_ptr.p, n, s);
    const long long ncand = read_scalar(cand_ptr.p + n, s);
    keys.alloc(ncand);
    cvals.alloc(ncand);
    launch_interp_expand(n, A.nnz, A.rp.p, A.ci.p, A.v.p, strong.p, A.cf.p, sum_strong.p, cnt_c.p, f2c.p, cand_ptr.p, keys.p,
                         cvals.p, s);
    segmented_sort_u64(keys.p, cand_ptr.p, n, s);
    oc.alloc(ncand);
    ov.alloc(ncand);
    pcnt.alloc(n);
    launch_interp_reduce(n, A.rp.p, A.ci.p, A.cf.p, f2c.p, cand_ptr.p, keys.p, cvals.p, prm_.trunc, prm_.pmax, oc.p,
                         ov.p, pcnt.p, s);
    A.prp.alloc(n + 1);
    scan_i32_to_i32(pcnt.p, A.prp.p, n, s);
    A.pnnz = read_scalar(A.prp.p + n, s);
    A.pci.alloc(A.pnnz);
    A.pv.alloc(A.pnnz);
    launch_interp_gather(n, cand_ptr.p, oc.p, ov.p, A.prp.p, A.pci.p, A.pv.p, s);
    A.nc = nc;
    // restriction R = P^T (rows sorted by fine index)
    colcnt.alloc(nc);
    launch_fill_i32(colcnt.p, 0, nc, s);
    launch_col_count(A.pnnz, A.pci.p, colcnt.p, s);
    rrp64.alloc(nc + 1);
    scan_i32_to_i64(colcnt.p, rrp64.p, nc, s);
    A.rrp.alloc(nc + 1);
    scan_i32_to_i32(colcnt.p, A.rrp.p, nc, s);
    cursor.alloc(nc);
    launch_fill_i32(cursor.p, 0, nc, s);
    {
      DevBuf<unsigned long long> tkeys;
      tkeys.alloc(A.pnnz);
      launch_transpose_fill(n, A.prp.p, A.pci.p, A.rrp.p, cursor.p, tkeys.p, s);
      segmented_sort_u64(tkeys.p, rrp64.p, nc, s);
      A.rci.alloc(A.pnnz);
      A.rv.alloc(A.pnnz);
      launch_transpose_finish(A.pnnz, tkeys.p, A.pv.p, A.rci.p, A.rv.p, s);
      sync_before_release(s);
    }
    // Galerkin product: AP = A*P, Ac = R*AP, both through sorted product lists
    Level nxt;
    auto spgemm = [&](int nrows, long long lnnz, const int* lrp, const int* lci, const int* mrp, const int* mci, DevBuf<int>& crp,
                      DevBuf<int>& cci, int& cnnz, DevBuf<long long>& seg, DevBuf<int>& pe, DevBuf<int>& pf,
                      int& group) {
      DevBuf<int> pc, uc;
      DevBuf<long long> pptr;
      DevBuf<unsigned long long> pk;
      pc.alloc(nrows);
      launch_spgemm_count(nrows, lrp, lci, mrp, pc.p, s);
      pptr.alloc(nrows + 1);
      scan_i32_to_i64(pc.p, pptr.p, nrows, s);
      const long long tot = read_scalar(pptr.p + nrows, s);
      IMPES_REQUIRE(tot < (1LL << 40), "amg setup: product list too large");
      pk.alloc(tot);
      DevBuf<int> ex_e, ex_f;
      ex_e.alloc(tot);
      ex_f.alloc(tot);
      launch_spgemm_expand(nrows, lnnz, tot, lrp, lci, mrp, mci, pptr.p, pk.p, ex_e.p, ex_f.p, s);
      segmented_sort_u64(pk.p, pptr.p, nrows, s);
      uc.alloc(nrows);
      launch_spgemm_unique_count(nrows, pptr.p, pk.p, uc.p, s);
      crp.alloc(nrows + 1);
      scan_i32_to_i32(uc.p, crp.p, nrows, s);
      cnnz = read_scalar(crp.p + nrows, s);
      cci.alloc(cnnz);
      seg.alloc(cnnz + 1);
      pe.alloc(tot);
      pf.alloc(tot);
      launch_spgemm_finalize(nrows, pptr.p, pk.p, crp.p, ex_e.p, ex_f.p, cci.p, seg.p, pe.p, pf.p, tot, cnnz, s);
      sync_before_release(s);
      group = spgemm_group_for(tot, cnnz);
    };
    spgemm(n, A.nnz, A.rp.p, A.ci.p, A.prp.p, A.pci.p, A.aprp, A.apci, A.apnnz, A.ap_seg, A.ap_pe, A.ap_pf, A.ap_group);
    A.apv.alloc(A.apnnz);
    launch_spgemm_numeric(A.apnnz, A.ap_seg.p, A.ap_pe.p, A.ap_pf.p, A.v.p, A.pv.p, A.apv.p, A.ap_group, s);
    nxt.n = nc;
    spgemm(nc, A.pnnz, A.rrp.p, A.rci.p, A.aprp.p, A.apci.p, nxt.rp, nxt.ci, nxt.nnz, A.ac_seg, A.ac_pe, A.ac_pf,
           A.ac_group);
    nxt.v.alloc(nxt.nnz);
    launch_spgemm_numeric(nxt.nnz, A.ac_seg.p, A.ac_pe.p, A.ac_pf.p, A.rv.p, A.apv.p, nxt.v.p, A.ac_group, s);
    A.pgroup = csr_group_for(n, A.pnnz);
    A.rgroup = csr_group_for(nc, A.pnnz);
    if (trace) IMPES_HIP_CHECK(hipStreamSynchronize(s));
    else sync_before_release(s);
    if (trace) {
      const double now = wall();
      std::fprintf(stderr, "[amg] level %d n=%d nnz=%d -> nc=%d p_nnz=%d ac_nnz=%d  %.3f s\n", lvl, n, nnz, nc,
                   A.pnnz, nxt.nnz, now - tl);
      tl = now;
    }
    levels_.push_back(std::move(nxt));
  }
  if (!levels_.back().coarsest) levels_.back().coarsest = true;
  for (int l = 0; l < (int)levels_.size(); ++l) finish_level_values(l, s);
}

void Hierarchy::finish_level_values(int l, hipStream_t s) {
  Level& L = levels_[l];
  L.group = csr_group_for(L.n, L.nnz);
  L.ir64.alloc(L.n);
  L.ir32.alloc(L.n);
  launch_diag_relax(L.n, L.rp.p, L.ci.p, L.v.p, nullptr, L.ir64.p, L.ir32.p, s);
  if (L.coarsest) build_coarsest(l, s);
}

void Hierarchy::build_coarsest(int l, hipStream_t s) {
  Level& L = levels_[l];
  L.dense = L.n <= prm_.dense_max;
  // a single-level hierarchy is solved without the pin: alone it would freeze DOF 0
  L.pin_mode = levels_.size() > 1 ? prm_.coarse_pin : kPinNever;
  L.pin.alloc(2);  // [0] pin flag, [1] inverse produced by the SPD sweep
  launch_pin_decide(L.n, L.rp.p, L.ci.p, L.v.p, L.pin_mode, L.pin.p, s);
  if (!L.dense) return;
  // above 160 rows the blocked matrix-core inverse is faster than the one-workgroup
  // sweep (a 508-row coarsest level: 0.3 ms per refresh against 85 ms)
  L.mfma = L.n > kDenseSweepMax;
  if (L.mfma) {
    const long long np = dense_pad(L.n);
    L.inv.alloc(np * np);
    L.gj_w.alloc(np * np);
    L.gj_p.alloc(2 * 64 * 64);
    L.gj_bad.alloc(1);
    dense_inverse(L, s);
    if (read_scalar(L.gj_bad.p, s) == 0) return;
    // a non-positive pivot: not SPD; up to 512 rows the pivoted elimination below
    IMPES_REQUIRE(L.n <= 512, "amg setup: coarsest operator above 512 rows is not symmetric positive definite "
                              "(lower dense_max)");
    L.mfma = false;
    L.inv.release();
    L.gj_w.release();
    L.gj_p.release();
    L.gj_bad.release();
  }
  L.aug.alloc(2LL * L.n * L.n);
  launch_csr_to_dense(L.n, L.rp.p, L.ci.p, L.v.p, L.pin.p, L.aug.p, s);
  launch_gauss_jordan_inverse(L.n, L.aug.p, L.pin.p + 1, s);
}

void Hierarchy::dense_inverse(Level& L, hipStream_t s) {
  launch_fill_i32(L.gj_bad.p, 0, 1, s);
  launch_dense_inverse_mfma(L.n, L.rp.p, L.ci.p, L.v.p, L.pin.p, L.inv.p, L.gj_w.p, L.gj_p.p, L.gj_bad.p, s);
}

void Hierarchy::build_pad4(Level& L, hipStream_t s) {
  DevBuf<int> cnt;
  cnt.alloc(L.n);
  L.rp4.alloc(L.n + 1);
  launch_csr4_count(L.n, L.rp.p, cnt.p, s);
  scan_i32_to_i32(cnt.p, L.rp4.p, L.n, s);
  const long long nnz4 = read_scalar(L.rp4.p + L.n, s);
  L.ci4.alloc(nnz4);
  L.v4.alloc(nnz4);
  launch_csr4_pattern(L.n, L.rp.p, L.ci.p, L.rp4.p, L.ci4.p, s);
  launch_csr4_values(L.n, L.rp.p, L.v.p, L.rp4.p, L.v4.p, s);
  L.g4 = csr4_group_for(L.n, nnz4);
  L.pad4 = true;
  // 16-bit column offsets from the row where every entry is within 2^15 of it
  DevBuf<int> maxabs;
  maxabs.alloc(1);
  launch_fill_i32(maxabs.p, 0, 1, s);
  L.co4.alloc(nnz4);
  launch_csr4_offsets(L.n, L.rp4.p, L.ci4.p, L.co4.p, maxabs.p, s);
  if (read_scalar(maxabs.p, s) <= 32767) L.ci4.release();  // 8.97 -> 8.88 ms per step at 1M cells
  else L.co4.release();
  // wave tiles of whole rows (the pattern is fixed after setup)
  std::vector<int> hrp(L.n + 1), wrow, wq;
  IMPES_HIP_CHECK(hipMemcpyAsync(hrp.data(), L.rp4.p, sizeof(int) * (L.n + 1), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  L.nw4 = csr4_wave_tiles(L.n, hrp.data(), wrow, wq);
  if (L.nw4 > 0) {
    L.wrow4.alloc(L.nw4 + 1);
    L.wq4.alloc(L.nw4 + 1);
    IMPES_HIP_CHECK(hipMemcpyAsync(L.wrow4.p, wrow.data(), sizeof(int) * (L.nw4 + 1), hipMemcpyHostToDevice, s));
    IMPES_HIP_CHECK(hipMemcpyAsync(L.wq4.p, wq.data(), sizeof(int) * (L.nw4 + 1), hipMemcpyHostToDevice, s));
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void Hierarchy::build_pad4_transfer(int rows, int nnz, const int* rp, const int* ci, const double* v, Pad4& T,
                                    hipStream_t s) {
  T = Pad4{};
  if (rows <= 0) return;
  DevBuf<int> cnt;
  cnt.alloc(rows);
  T.rp4.alloc(rows + 1LL);
  launch_csr4_count(rows, rp, cnt.p, s);
  scan_i32_to_i32(cnt.p, T.rp4.p, rows, s);
  std::vector<int> hrp(rows + 1LL), wrow, wq;
  IMPES_HIP_CHECK(hipMemcpyAsync(hrp.data(), T.rp4.p, sizeof(int) * (rows + 1LL), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  const int nw = csr4_wave_tiles(rows, hrp.data(), wrow, wq);
  if (nw == 0) {  // a row longer than a tile: the CSR transfer kernels stay
    T.rp4.release();
    return;
  }
  const long long nnz4 = hrp[rows];
  T.ci4.alloc(nnz4);
  T.v4.alloc(nnz4);
  launch_csr4_pattern(rows, rp, ci, T.rp4.p, T.ci4.p, s, false);
  launch_csr4_values(rows, rp, v, T.rp4.p, T.v4.p, s);
  T.wrow.alloc(nw + 1LL);
  T.wq.alloc(nw + 1LL);
  IMPES_HIP_CHECK(hipMemcpyAsync(T.wrow.p, wrow.data(), sizeof(int) * (nw + 1LL), hipMemcpyHostToDevice, s));
  IMPES_HIP_CHECK(hipMemcpyAsync(T.wq.p, wq.data(), sizeof(int) * (nw + 1LL), hipMemcpyHostToDevice, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  T.n = rows;
  T.nnz = nnz;
  T.nw = nw;
}

void Hierarchy::refresh_f32(hipStream_t s) {
  for (int l = 0; l < (int)levels_.size(); ++l) {
    Level& L = levels_[l];
    L.pad4 = false;
    if (!(l == 0 && stencil_)) build_pad4(L, s);
    if (!L.pad4) {  // the unpadded FP32 values serve the levels without the padded form
      L.v32.alloc(L.nnz);
      launch_f64_to_f32(L.v32.p, L.v.p, L.nnz, s);
    } else {
      L.v32.release();
    }
    if (!L.coarsest) {
      L.pv32.alloc(L.pnnz);
      L.rv32.alloc(L.pnnz);
      launch_f64_to_f32(L.pv32.p, L.pv.p, L.pnnz, s);
      launch_f64_to_f32(L.rv32.p, L.rv.p, L.pnnz, s);
      // restriction rows (~5 entries) in the padded tiled form: level-0/1
      // restrictions 11.8 / 8.3 -> 9.8 / 6.4 us at 1M cells; interpolation rows
      // (~2 entries, one lane per row) stay CSR: padded they took 10.1 -> 12.0 us
      build_pad4_transfer(L.nc, L.pnnz, L.rrp.p, L.rci.p, L.rv.p, L.r4, s);
    }
  }
}

void Hierarchy::refresh_stencil(hipStream_t s) {
  const long long n = (long long)nx_ * ny_ * nz_;
  sd_.alloc(n); sx_.alloc(n); sy_.alloc(n); sz_.alloc(n);
  sd32_.alloc(n); sx32_.alloc(n); sy32_.alloc(n); sz32_.alloc(n);
  k_st7_scale<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(nx_, ny_, nz_, st_d_.p, st_x_.p, st_y_.p, st_z_.p, dh_.p,
                                                         sd_.p, sx_.p, sy_.p, sz_.p, sd32_.p, sx32_.p, sy32_.p,
                                                         sz32_.p);
  IMPES_LAUNCH_CHECK();
}

void Hierarchy::alloc_vectors() {
  for (auto& L : levels_) {
    L.x64.alloc(L.n); L.b64.alloc(L.n); L.r64.alloc(L.n); L.t64.alloc(L.n);
    L.x32.alloc(L.n); L.b32.alloc(L.n); L.r32.alloc(L.n); L.t32.alloc(L.n);
  }
}

// ---------------------------------------------------------------- update
void Hierarchy::update_csr_values(const double* v, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "amg update before setup");
  if (v != v0_unscaled_.p) d2d(v0_unscaled_.p, v, levels_[0].nnz, s);
  equilibrate(v0_unscaled_.p, false, s);
  update_levels(false, s);
}

void Hierarchy::update_levels(bool fine_done, hipStream_t s) {
  for (int l = 0; l + 1 < (int)levels_.size(); ++l) {
    Level& A = levels_[l];
    if (A.coarsest) break;
    Level& C = levels_[l + 1];
    launch_spgemm_numeric(A.apnnz, A.ap_seg.p, A.ap_pe.p, A.ap_pf.p, A.v.p, A.pv.p, A.apv.p, A.ap_group, s);
    launch_spgemm_numeric(C.nnz, A.ac_seg.p, A.ac_pe.p, A.ac_pf.p, A.rv.p, A.apv.p, C.v.p, A.ac_group, s);
  }
  for (int l = 0; l < (int)levels_.size(); ++l) {
    Level& L = levels_[l];
    if (L.coarsest && L.dense) {
      // the dense solve reads only the FP64 values: no FP32 copies (the weights stay
      // current for the level's exported data)
      launch_diag_relax(L.n, L.rp.p, L.ci.p, L.v.p, nullptr, L.ir64.p, L.ir32.p, s);
    } else if (!(fine_done && l == 0)) {
      if (L.pad4) {
        launch_diag_relax_pad4(L.n, L.rp.p, L.ci.p, L.v.p, L.rp4.p, L.v4.p, L.ir64.p, L.ir32.p, s);
      } else {
        launch_diag_relax(L.n, L.rp.p, L.ci.p, L.v.p, nullptr, L.ir64.p, L.ir32.p, s);
        launch_f64_to_f32(L.v32.p, L.v.p, L.nnz, s);
      }
    }
    if (L.coarsest) {
      launch_pin_decide(L.n, L.rp.p, L.ci.p, L.v.p, L.pin_mode, L.pin.p, s);
      if (L.dense && L.mfma) {
        dense_inverse(L, s);
      } else if (L.dense) {
        launch_csr_to_dense(L.n, L.rp.p, L.ci.p, L.v.p, L.pin.p, L.aug.p, s);
        launch_gauss_jordan_inverse(L.n, L.aug.p, L.pin.p + 1, s);
      }
    }
  }
  if (stencil_ && !fine_done) refresh_stencil(s);
}

void Hierarchy::update_stencil(const double* d, const double* ox, const double* oy, const double* oz, hipStream_t s) {
  IMPES_REQUIRE(stencil_ && !levels_.empty(), "amg update_stencil on a CSR hierarchy");
  const long long n = (long long)nx_ * ny_ * nz_;
  Level& L0 = levels_[0];
  // Level 0 runs on the stencil kernels: its relaxation weights and equilibrated
  // stencil come from the fused pass, and its CSR values only feed the Galerkin
  // products (the FP32 copy of level-0 CSR values is not read in this mode).
  const bool scalar = prm_.equil_mode == kEquilScalar;
  if (!scalar) equil_factors(d, false, s);
  St7FineUpdate u{nx_, ny_, nz_, d, ox, oy, oz, scalar ? nullptr : dh_.p, scalar_, L0.rp.p,
                  st_d_.p, st_x_.p, st_y_.p, st_z_.p, diag0_.p, L0.v.p, L0.ir64.p, L0.ir32.p,
                  sd_.p, sx_.p, sy_.p, sz_.p, sd32_.p, sx32_.p, sy32_.p, sz32_.p, dh_.p, sin_.p, sout_.p};
  const int blocks = (int)std::min<long long>((n + kFineRows - 1) / kFineRows, 1 << 16);
  if (scalar) klaunch(k_st7_fine_update<true>, dim3(blocks), dim3(kFineRows), 0, s, u);
  else klaunch(k_st7_fine_update<false>, dim3(blocks), dim3(kFineRows), 0, s, u);
  IMPES_LAUNCH_CHECK();
  update_levels(true, s);
}

bool Hierarchy::coarse_pinned(hipStream_t s) const {
  IMPES_REQUIRE(!levels_.empty(), "amg hierarchy not set up");
  int h = 0;
  IMPES_HIP_CHECK(hipMemcpyAsync(&h, levels_.back().pin.p, sizeof(int), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  return h != 0;
}

LevelInfo Hierarchy::info(int l) const {
  const Level& L = levels_.at(l);
  return LevelInfo{L.n, L.nnz, L.coarsest ? 0 : L.pnnz, L.coarsest ? 1 : 0, L.dense ? 1 : 0};
}

// ---------------------------------------------------------------- V-cycle
template <class T>
void Hierarchy::smooth(int l, const T* b, T*& cur, T* other, int sweeps, bool& zero, hipStream_t s) {
  Level& L = levels_[l];
  const bool st = (l == 0 && stencil_);
  for (int it = 0; it < sweeps; ++it) {
    if (zero && st && it + 1 < sweeps) {
      // the sweep from zero and the next one as one stencil pass
      if constexpr (sizeof(T) == 8) {
        Stencil7<double> S{nx_, ny_, nz_, sd_.p, sx_.p, sy_.p, sz_.p};
        launch_st7_jacobi_z0_f64(S, (const double*)b, L.ir64.p, prm_.omega, (double*)other, s);
      } else {
        Stencil7<float> S{nx_, ny_, nz_, sd32_.p, sx32_.p, sy32_.p, sz32_.p};
        launch_st7_jacobi_z0_f32(S, (const float*)b, L.ir32.p, (float)prm_.omega, (float*)other, s);
      }
      zero = false;
      std::swap(cur, other);
      ++it;
      continue;
    }
    if (zero) {
      Sel<T>::jac0(b, Sel<T>::ir(L), prm_.omega, cur, L.n, s);
      zero = false;
      continue;
    }
    if (st) {
      if constexpr (sizeof(T) == 8) {
        Stencil7<double> S{nx_, ny_, nz_, sd_.p, sx_.p, sy_.p, sz_.p};
        launch_st7_jacobi_f64(S, (const double*)cur, (const double*)b, L.ir64.p, prm_.omega, (double*)other, s);
      } else {
        Stencil7<float> S{nx_, ny_, nz_, sd32_.p, sx32_.p, sy32_.p, sz32_.p};
        launch_st7_jacobi_f32(S, (const float*)cur, (const float*)b, L.ir32.p, (float)prm_.omega, (float*)other, s);
      }
    } else {
      Sel<T>::jacL(L, cur, b, prm_.omega, other, s);
    }
    std::swap(cur, other);
  }
}

template <class T>
void Hierarchy::coarse_solve(int l, const T* b, T* x, hipStream_t s) {
  Level& L = levels_[l];
  if (L.dense && L.mfma) {
    Sel<T>::dense_mv(L.n, L.pin.p, L.inv.p, b, x, s);
    return;
  }
  if (L.dense) {
    Sel<T>::dense(L.n, L.pin.p, L.aug.p, b, x, s);
    return;
  }
  // reference: 50 L1-Jacobi sweeps from zero on the unpinned matrix with b_0 = 0
  T* bl = Sel<T>::r(L);
  Sel<T>::copy(bl, b, L.n, s);
  Sel<T>::pin0(L.pin.p, bl, s);
  T* cur = x;
  T* other = Sel<T>::t(L);
  bool zero = true;
  smooth<T>(l, bl, cur, other, prm_.coarse_sweeps, zero, s);
  if (cur != x) Sel<T>::copy(x, cur, L.n, s);
}

int Hierarchy::mark(hipStream_t s) {
  if (!clock_) return -1;
  Clock& c = *clock_;
  if (c.used == (int)c.ev.size()) {
    hipEvent_t e = nullptr;
    IMPES_HIP_CHECK(hipEventCreate(&e));
    c.ev.push_back(e);
  }
  IMPES_HIP_CHECK(hipEventRecord(c.ev[c.used], s));
  return c.used++;
}

double Hierarchy::level_bytes(int l, int precision) const {
  const Level& L = levels_.at(l);
  const double vb = precision == 32 ? 4.0 : 8.0;
  const double n = L.n;
  // a CSR pass over `rows` rows with `nnz` entries gathering a vector of `cols`
  // entries and streaming `vecs` more row vectors
  auto csr = [&](double rows, double nnz, double cols, double vecs) {
    return (rows + 1) * 4.0 + nnz * (4.0 + vb) + cols * vb + rows * vb * vecs;
  };
  const bool st = l == 0 && stencil_;
  auto sweep = [&](double vecs) {  // L1-Jacobi (vecs 3: b, inv_relax, out) or residual (vecs 2: b, out)
    return st ? n * vb * (4.0 + 1.0 + vecs) : csr(n, L.nnz, n, vecs);
  };
  if (L.coarsest) {
    if (L.dense) return n * n * 8.0 + 2.0 * n * vb;
    return 2.0 * n * vb + n * vb * 3.0 + (prm_.coarse_sweeps - 1) * sweep(3.0);
  }
  const Level& C = levels_[l + 1];
  const bool first_done = l > 0 && prm_.pre_smooth > 0;  // fused into the parent's restriction
  double b = 0.0;
  int pre = prm_.pre_smooth - (first_done ? 1 : 0);
  if (st && !first_done && pre > 1) {
    b += n * vb * (4.0 + 3.0);  // the two sweeps from zero in one stencil pass: coefficients, b, inv_relax, out
    pre -= 2;
  } else if (!first_done && pre > 0) {
    b += 3.0 * n * vb;  // first sweep from zero: b, inv_relax -> x
    --pre;
  }
  b += pre * sweep(3.0);
  b += (prm_.pre_smooth == 0) ? 2.0 * n * vb : sweep(2.0);  // residual (a copy without pre-smoothing)
  const bool fused = !C.coarsest && prm_.pre_smooth > 0;
  b += csr(C.n, L.pnnz, n, fused ? 3.0 : 1.0);  // restriction (+ the coarse first sweep: inv_relax, b_c, x_c)
  b += csr(n, L.pnnz, C.n, 2.0);                 // prolongation: x + P e_c
  b += prm_.post_smooth * sweep(3.0);
  if (prm_.post_smooth % 2) b += 2.0 * n * vb;  // the final copy back into x
  return b;
}

void Hierarchy::level_times(int precision, int cycles, double* ms, double* cycle_ms, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "level_times before setup");
  IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
  IMPES_REQUIRE(cycles >= 1, "cycles must be >= 1");
  Clock c;
  struct Done {
    Hierarchy* h;
    Clock& c;
    ~Done() {
      h->clock_ = nullptr;
      for (auto e : c.ev) (void)hipEventDestroy(e);
    }
  } done{this, c};
  Level& L0 = levels_[0];
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  clock_ = &c;
  for (int i = 0; i < cycles; ++i) {
    if (precision == 32) vcycle_level<float>(0, L0.b32.p, L0.x32.p, true, s);
    else vcycle_level<double>(0, L0.b64.p, L0.x64.p, true, s);
  }
  clock_ = nullptr;
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<double> acc(levels_.size(), 0.0);
  for (const auto& g : c.seg) {
    float t = 0.f;
    IMPES_HIP_CHECK(hipEventElapsedTime(&t, c.ev[g[1]], c.ev[g[2]]));
    acc[g[0]] += t;
  }
  for (size_t l = 0; l < acc.size(); ++l) ms[l] = acc[l] / cycles;
  if (cycle_ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    IMPES_HIP_CHECK(hipEventCreate(&e0));
    IMPES_HIP_CHECK(hipEventCreate(&e1));
    struct Ev {
      hipEvent_t a, b;
      ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    vcycle_only(precision, s);  // captures the graph when needed
    IMPES_HIP_CHECK(hipEventRecord(e0, s));
    for (int i = 0; i < cycles; ++i) vcycle_only(precision, s);
    IMPES_HIP_CHECK(hipEventRecord(e1, s));
    IMPES_HIP_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    IMPES_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    *cycle_ms = t / cycles;
  }
}

template <class T>
void Hierarchy::vcycle_level(int l, const T* b, T* x, bool x_zero, hipStream_t s, bool first_sweep_done) {
  Level& L = levels_[l];
  const int m0 = mark(s);
  if (L.coarsest) {
    coarse_solve<T>(l, b, x, s);
    segment(l, m0, mark(s));
    return;
  }
  T* cur = x;
  T* other = Sel<T>::t(L);
  bool zero = x_zero && !first_sweep_done;
  smooth<T>(l, b, cur, other, prm_.pre_smooth - (first_sweep_done ? 1 : 0), zero, s);
  T* r = Sel<T>::r(L);
  if (zero) {
    Sel<T>::copy(r, b, L.n, s);  // no pre-smoothing: r = b - A*0
  } else if (l == 0 && stencil_) {
    if constexpr (sizeof(T) == 8) {
      Stencil7<double> S{nx_, ny_, nz_, sd_.p, sx_.p, sy_.p, sz_.p};
      launch_st7_residual_f64(S, (const double*)cur, (const double*)b, (double*)r, s);
    } else {
      Stencil7<float> S{nx_, ny_, nz_, sd32_.p, sx32_.p, sy32_.p, sz32_.p};
      launch_st7_residual_f32(S, (const float*)cur, (const float*)b, (float*)r, s);
    }
  } else {
    Sel<T>::resL(L, cur, b, r, s);
  }
  Level& C = levels_[l + 1];
  const bool fused = !C.coarsest && prm_.pre_smooth > 0;
  if (fused) {
    // restriction fused with the coarse level's first sweep from zero
    Sel<T>::restrict_j0(L, r, Sel<T>::ir(C), prm_.omega, Sel<T>::b(C), Sel<T>::x(C), s);
  } else {
    Sel<T>::restrict_to(L, r, Sel<T>::b(C), s);
  }
  segment(l, m0, mark(s));
  vcycle_level<T>(l + 1, Sel<T>::b(C), Sel<T>::x(C), true, s, fused);
  const int m1 = mark(s);
  // Post-smoothing alternates between x and t: prolongate into the buffer from
  // which post_smooth sweeps end in x, so the cycle needs no final copy.
  T* target = (prm_.post_smooth % 2 == 0) ? x : Sel<T>::t(L);
  if (zero) {
    // no pre-smoothing happened: x = P e_c
    Sel<T>::prolong(L, Sel<T>::x(C), target, s);
    zero = false;
  } else if (target == cur) {
    Sel<T>::prolong_add(L, Sel<T>::x(C), cur, s);
  } else {
    Sel<T>::prolong_add_to(L, Sel<T>::x(C), cur, target, s);
  }
  cur = target;
  other = (cur == x) ? Sel<T>::t(L) : x;
  smooth<T>(l, b, cur, other, prm_.post_smooth, zero, s);
  if (cur != x) Sel<T>::copy(x, cur, L.n, s);
  segment(l, m1, mark(s));
}

namespace {
// swaps the smoothing counts for one eager cycle (restored on every exit path)
struct SmoothSwap {
  AmgParams& p;
  bool on;
  SmoothSwap(AmgParams& prm, bool swap) : p(prm), on(swap) { if (on) std::swap(p.pre_smooth, p.post_smooth); }
  ~SmoothSwap() { if (on) std::swap(p.pre_smooth, p.post_smooth); }
};
}  // namespace

void Hierarchy::vcycle64(const double* b, double* x, bool x_zero, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "vcycle before setup");
  SmoothSwap sw(prm_, transposed_);
  vcycle_level<double>(0, b, x, x_zero, s);
}

void Hierarchy::vcycle32(const float* b, float* x, bool x_zero, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "vcycle before setup");
  SmoothSwap sw(prm_, transposed_);
  vcycle_level<float>(0, b, x, x_zero, s);
}

int Hierarchy::cycles(int precision, const double* b, double* x, bool use_x0, int min_cycles, double tol,
                      int max_cycles, double* rel, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "cycles before setup");
  const long long n = levels_[0].n;
  cyc_b_.reserve(n); cyc_x_.reserve(n); cyc_r_.reserve(n);
  cyc_s_.reserve(kRedBlocks + 4);
  CycleWork w{cyc_b_.p, cyc_x_.p, cyc_r_.p, cyc_s_.p, cyc_s_.p + kRedBlocks};
  return cycles_impl(w, precision, b, x, use_x0, min_cycles, tol, max_cycles, rel, s);
}

int Hierarchy::cycles_in(double* ws, int precision, const double* b, double* x, bool use_x0, int min_cycles,
                         double tol, int max_cycles, double* rel, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "cycles before setup");
  IMPES_REQUIRE(ws && (reinterpret_cast<uintptr_t>(ws) & 255) == 0, "cycles: workspace must be 256-byte aligned");
  const long long n = levels_[0].n;
  // the n-vectors start on 256-byte boundaries (n rounded up to 32 doubles)
  const long long ld = (n + 31) / 32 * 32;
  CycleWork w{ws, ws + ld, ws + 2 * ld, ws + 3 * ld, ws + 3 * ld + kRedBlocks};
  return cycles_impl(w, precision, b, x, use_x0, min_cycles, tol, max_cycles, rel, s);
}

int Hierarchy::cycles_impl(const CycleWork& w, int precision, const double* b, double* x, bool use_x0,
                           int min_cycles, double tol, int max_cycles, double* rel, hipStream_t s) {
  Level& L0 = levels_[0];
  const long long n = L0.n;
  double hv[2];
  launch_scale_f64(w.b, dh_.p, b, n, s);  // bs = dh .* b
  if (use_x0) launch_div_f64(w.x, x, dh_.p, n, s);  // xs = x0 ./ dh
  launch_dot_f64(w.b, w.b, n, w.sc, w.part, s);
  IMPES_HIP_CHECK(hipMemcpyAsync(hv, w.sc, sizeof(double), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(hv[0]);
  bool zero = !use_x0;
  int c = 0;
  double r = INFINITY;
  while (c < max_cycles) {
    if (precision == 32) {
      launch_f64_to_f32(L0.b32.p, w.b, n, s);
      if (!zero) launch_f64_to_f32(L0.x32.p, w.x, n, s);
      vcycle_level<float>(0, L0.b32.p, L0.x32.p, zero, s);
      launch_f32_to_f64(w.x, L0.x32.p, n, s);
    } else {
      vcycle_level<double>(0, w.b, w.x, zero, s);
    }
    zero = false;
    ++c;
    if (stencil_) {
      Stencil7<double> S{nx_, ny_, nz_, sd_.p, sx_.p, sy_.p, sz_.p};
      launch_st7_residual_f64(S, w.x, w.b, w.r, s);
    } else {
      launch_residual_f64(Sel<double>::A(L0), w.x, w.b, w.r, L0.group, s);
    }
    launch_dot_f64(w.r, w.r, n, w.sc + 1, w.part, s);
    IMPES_HIP_CHECK(hipMemcpyAsync(hv + 1, w.sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
    r = std::sqrt(hv[1]) / (bnorm + 1e-30);
    if (c >= min_cycles && r < tol) break;
  }
  launch_scale_f64(x, dh_.p, w.x, n, s);
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  if (rel) *rel = r;
  return c;
}

void* Hierarchy::vcycle_in(int precision) {
  IMPES_REQUIRE(!levels_.empty(), "preconditioner before setup");
  return precision == 32 ? (void*)levels_[0].b32.p : (void*)levels_[0].b64.p;
}

const void* Hierarchy::vcycle_out(int precision) const {
  IMPES_REQUIRE(!levels_.empty(), "preconditioner before setup");
  return precision == 32 ? (const void*)levels_[0].x32.p : (const void*)levels_[0].x64.p;
}

void Hierarchy::vcycle_only(int precision, hipStream_t s) {
  IMPES_REQUIRE(!levels_.empty(), "preconditioner before setup");
  Level& L0 = levels_[0];
  const bool replay = graphs_ && !ktimer_enabled();
  const int gi = precision == 32 ? 0 : 1;
  if (replay && !vexec_[gi]) capture_vcycle(precision);
  if (replay) IMPES_HIP_CHECK(hipGraphLaunch(vexec_[gi], s));
  else if (precision == 32) vcycle_level<float>(0, L0.b32.p, L0.x32.p, true, s);
  else vcycle_level<double>(0, L0.b64.p, L0.x64.p, true, s);
}

void Hierarchy::precondition(int precision, const double* v, double* z, hipStream_t s) {
  Level& L0 = levels_[0];
  const long long n = L0.n;
  if (precision == 32) launch_scale_f64_to_f32(L0.b32.p, sin_.p, v, n, s);
  else launch_scale_f64(L0.b64.p, sin_.p, v, n, s);
  vcycle_only(precision, s);
  if (precision == 32) launch_scale_f32_to_f64(z, sout_.p, L0.x32.p, n, s);
  else launch_scale_f64(z, sout_.p, L0.x64.p, n, s);
}

}  // namespace impes
