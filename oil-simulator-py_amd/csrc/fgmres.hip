// FP64 restarted FGMRES(m) with CGS2 orthogonalisation. The Hessenberg matrix,
// Givens rotations and the residual estimate live on the device; the host reads
// one scalar per Arnoldi step (written by the device into pinned host memory) to
// decide termination.
//
// One Arnoldi step with the AMG preconditioner is, besides the V-cycle:
//   st7/CSR operator apply (fused with the V-cycle's output scaling on stencils),
//   three Gram-Schmidt passes (dots; update + dots; update + norm + rotation),
//   and the normalisation of the new basis vector, fused with the V-cycle's
//   input scaling for the next step.
#include "fgmres.h"

#include <algorithm>
#include <cstdlib>

#include <cmath>

namespace impes {

namespace {

constexpr int kGsThreads = 256;
constexpr int kMaxFusedGs = 32;  // largest k of the register-resident Gram-Schmidt passes

struct Layout {
  int m;
  __host__ __device__ int H(int i, int j) const { return i + j * (m + 1); }
  __host__ __device__ int g() const { return (m + 1) * m; }
  __host__ __device__ int cs() const { return g() + m + 1; }
  __host__ __device__ int sn() const { return cs() + m; }
  __host__ __device__ int h() const { return sn() + m; }
  __host__ __device__ int h2() const { return h() + m + 1; }
  __host__ __device__ int y() const { return h2() + m + 1; }
  __host__ __device__ int sc() const { return y() + m; }  // 8 scalars
  __host__ __device__ int size() const { return sc() + 8; }
};

// sc[0] = beta = sqrt(sc[4]) ; g = beta e1
__global__ void k_init_cycle(Layout L, double* __restrict__ sm) {
  double* sc = sm + L.sc();
  const double beta = sqrt(sc[4]);
  sc[0] = (beta > 0.0) ? beta : 1.0;
  double* g = sm + L.g();
  for (int i = 0; i <= L.m; ++i) g[i] = 0.0;
  g[0] = beta;
}

// Column j of H from the two Gram-Schmidt passes (h + h2, ||w||^2 = ww), the
// previous rotations applied, a new rotation, the residual estimate |g[j+1]| in
// sc[2] (and in *est, pinned host memory) and the normalisation sc[3] = ||w||.
// Called by every thread of one block: the column and the rotations are staged in
// LDS, thread 0 runs the (sequential) rotation recurrence.
__device__ void arnoldi_column(Layout L, int j, double* __restrict__ sm, double ww, double* est) {
  __shared__ double col[kMaxMultiDot + 1], lc[kMaxMultiDot], ls[kMaxMultiDot];
  const double* h = sm + L.h();
  const double* h2 = sm + L.h2();
  double* cs = sm + L.cs();
  double* sn = sm + L.sn();
  for (int i = threadIdx.x; i <= j; i += blockDim.x) col[i] = h[i] + h2[i];
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    lc[i] = cs[i];
    ls[i] = sn[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* sc = sm + L.sc();
  double* g = sm + L.g();
  sc[1] = ww;
  const double hn = sqrt(fmax(ww, 0.0));
  col[j + 1] = hn;
  for (int i = 0; i < j; ++i) {
    const double a = col[i], b = col[i + 1];
    col[i] = lc[i] * a + ls[i] * b;
    col[i + 1] = -ls[i] * a + lc[i] * b;
  }
  const double a = col[j], b = col[j + 1];
  double c, s;
  if (b == 0.0) { c = 1.0; s = 0.0; }
  else if (fabs(b) > fabs(a)) { const double t = a / b; s = 1.0 / sqrt(1.0 + t * t); c = s * t; }
  else { const double t = b / a; c = 1.0 / sqrt(1.0 + t * t); s = c * t; }
  cs[j] = c;
  sn[j] = s;
  col[j] = c * a + s * b;
  col[j + 1] = 0.0;
  for (int i = 0; i <= j + 1; ++i) sm[L.H(i, j)] = col[i];
  g[j + 1] = -s * g[j];
  g[j] = c * g[j];
  sc[2] = fabs(g[j + 1]);
  sc[3] = (hn > 0.0) ? hn : 1.0;
  *est = sc[2];  // made visible to the host by the event recorded after this kernel
}

// Gram-Schmidt pass of Arnoldi step j over k = j + 1 basis vectors (k <= K). Each
// thread loads the k basis entries of its rows once into registers:
//   PASS 0: h = V^T w
//   PASS 1: w -= V h;  h2 = V^T w   (first update and second pass's dots, one sweep)
//   PASS 2: w -= V h2; ||w||^2
// Blocks leave partial sums in part (column c at part + c * gridDim.x), folded by
// k_fold (passes 0, 1) or k_arnoldi (pass 2). A fold inside the kernel by its last
// block was measured slower: the device-scope release every block then needs
// writes back its XCD's L2 (150-280 us per pass against ~25 us).
template <int K, int PASS>
__global__ __launch_bounds__(kGsThreads) void k_gs_pass(int k, double* __restrict__ w,
                                                        const double* __restrict__ V, long long n,
                                                        const double* __restrict__ hsrc, double* __restrict__ part) {
  constexpr bool UPDATE = PASS > 0, DOTS = PASS < 2;
  constexpr int NC = DOTS ? K : 1;
  constexpr int NW = kGsThreads / kWave;
  __shared__ double lds[NW * NC];
  double hc[K];
#pragma unroll
  for (int t = 0; t < K; ++t) hc[t] = (UPDATE && t < k) ? hsrc[t] : 0.0;
  double acc[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) acc[t] = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double vv[K];
#pragma unroll
    for (int t = 0; t < K; ++t) vv[t] = (t < k) ? V[(long long)t * n + i] : 0.0;
    double wi = w[i];
    if (UPDATE) {
#pragma unroll
      for (int t = 0; t < K; ++t)
        if (t < k) wi -= vv[t] * hc[t];
      w[i] = wi;
    }
    if (DOTS) {
#pragma unroll
      for (int t = 0; t < K; ++t) acc[t] += vv[t] * wi;
    } else {
      acc[0] += wi * wi;
    }
  }
  const int ncols = DOTS ? k : 1;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    if (t < ncols) {
      const double v = wave_sum(acc[t]);
      if (lane == 0) lds[wid * NC + t] = v;
    }
  }
  __syncthreads();
  const int tc = threadIdx.x;
  if (tc < ncols) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) v += lds[q * NC + tc];
    part[(long long)tc * gridDim.x + blockIdx.x] = v;
  }
}

// fold nb partials per column for k columns: out[c] = sum_b part[c*nb + b]
__global__ void k_fold_cols(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double lds[kGsThreads / kWave];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long long)c * nb + b];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) out[c] = s;
}

template <int PASS>
void launch_gs_pass(int k, double* w, const double* V, long long n, const double* h, double* part, hipStream_t s) {
  if (k <= 4)
    k_gs_pass<4, PASS><<<kMultiDotBlocks, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else if (k <= 8)
    k_gs_pass<8, PASS><<<kMultiDotBlocks, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else if (k <= 16)
    k_gs_pass<16, PASS><<<kMultiDotBlocks, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else
    k_gs_pass<kMaxFusedGs, PASS><<<kMultiDotBlocks, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  IMPES_LAUNCH_CHECK();
}

// Column j of H: folds the nb partial sums of ||w||^2, then arnoldi_column.
__global__ void k_arnoldi(Layout L, int j, double* __restrict__ sm, const double* __restrict__ part, int nb,
                          double* est) {
  __shared__ double lds[kGsThreads / kWave];
  double ww = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) ww += part[b];
  ww = block_sum(ww, lds);
  arnoldi_column(L, j, sm, ww, est);
}

// v = w / nrm[0] and the next V-cycle input b = a .* v (in the hierarchy's
// precision). v may alias w.
template <class TO>
__global__ void k_normalize_scale(double* v, const double* w, const double* __restrict__ nrm,
                                  const double* __restrict__ a, TO* __restrict__ b, long long n) {
  const double inv = 1.0 / nrm[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double vi = w[i] * inv;
    v[i] = vi;
    b[i] = (TO)(a[i] * vi);
  }
}

__global__ void k_backsolve(Layout L, int k, double* __restrict__ sm) {
  const double* g = sm + L.g();
  double* y = sm + L.y();
  for (int i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int c = i + 1; c < k; ++c) s -= sm[L.H(i, c)] * y[c];
    const double d = sm[L.H(i, i)];
    y[i] = (d != 0.0) ? s / d : 0.0;
  }
}

}  // namespace

constexpr double kDeferMargin = 4.0;
constexpr double kErrAccept = 0.5;   // accept an error estimate up to this fraction of etol
constexpr double kErrTarget = 0.25;  // residual targets aim the estimate at this fraction of etol

Fgmres::Fgmres() {
  const char* e = std::getenv("IMPES_FGMRES_LOOKAHEAD");
  lookahead_ = e && e[0] == '1';
  no_defer_ = e && e[0] == '0';
}

Fgmres::~Fgmres() {
  if (host_) (void)hipHostFree(host_);
  for (auto& e : est_ev_)
    if (e) (void)hipEventDestroy(e);
}

void Fgmres::ensure(long long n, int m) {
  if (!host_) {
    IMPES_HIP_CHECK(hipHostMalloc((void**)&host_, 8 * sizeof(double), hipHostMallocMapped));
    IMPES_HIP_CHECK(hipHostGetDevicePointer((void**)&host_dev_, host_, 0));
  }
  for (auto& e : est_ev_)
    if (!e) IMPES_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (n == n_ && m == m_) return;
  n_ = n;
  m_ = m;
  V_.alloc((size_t)(m + 1) * n);
  Z_.alloc((size_t)m * n);
  w_.alloc(n);
  bs_.alloc(n);
  Layout L{m};
  small_.alloc(L.size());
  work_.alloc((size_t)std::max(kRedBlocks, kMultiDotBlocks) * (m + 1));
}

void Fgmres::apply(const Operator& A, const double* x, double* y, hipStream_t s) {
  if (A.stencil) launch_st7_rowscaled_spmv_f64(A.st, A.dinv, x, y, s);
  else launch_csr_rowscaled_spmv_f64(A.csr, A.dinv, x, y, A.group, s);
}

void Fgmres::residual(const Operator& A, const double* x, const double* b, double* r, hipStream_t s) {
  if (A.stencil) launch_st7_rowscaled_residual_f64(A.st, A.dinv, x, b, r, s);
  else launch_csr_rowscaled_residual_f64(A.csr, A.dinv, x, b, r, A.group, s);
}

void Fgmres::solve(const Operator& A, Hierarchy* M, int precision, const double* b, double* x, bool use_x0,
                   double rtol, int restart, int maxit, SolveStats& st, hipStream_t s) {
  const long long n = A.n;
  IMPES_REQUIRE(n >= 1, "fgmres: empty system");
  IMPES_REQUIRE(restart >= 1 && restart < kMaxMultiDot, "fgmres: restart out of range");
  IMPES_REQUIRE(M == nullptr || M->n0() == n, "fgmres: preconditioner size mismatch");
  IMPES_REQUIRE(precision == 32 || precision == 64, "fgmres: precision must be 32 or 64");
  ensure(n, restart);
  const int m = restart;
  Layout L{m};
  double* sm = small_.p;
  double* sc = sm + L.sc();
  auto read = [&](const double* dev) {
    IMPES_HIP_CHECK(hipMemcpyAsync(host_, dev, sizeof(double), hipMemcpyDeviceToHost, s));
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
    return host_[0];
  };
  st = SolveStats{};
  st.precision = precision;
  // ||D^-1 b||
  launch_scale_f64(bs_.p, A.dinv, b, n, s);
  launch_dot_f64(bs_.p, bs_.p, n, sc + 5, work_.p, s);
  const double bnorm = std::sqrt(read(sc + 5));
  st.bnorm = bnorm;
  if (!use_x0) launch_fill_f64(x, 0.0, n, s);
  if (bnorm == 0.0) {
    launch_fill_f64(x, 0.0, n, s);
    st.converged = 1;
    return;
  }
  const double tol = rtol * bnorm;
  double* V = V_.p;
  double* Z = Z_.p;
  // v = w / nrm, and with a preconditioner also its scaled V-cycle input
  auto normalize = [&](double* v, const double* w, const double* nrm) {
    if (!M) {
      launch_scale_by_inv_f64(v, w, nrm, n, s);
      return;
    }
    const int grid = grid_for(n, kGsThreads);
    if (precision == 32)
      k_normalize_scale<float><<<grid, kGsThreads, 0, s>>>(v, w, nrm, M->in_scale(), (float*)M->vcycle_in(32), n);
    else
      k_normalize_scale<double><<<grid, kGsThreads, 0, s>>>(v, w, nrm, M->in_scale(), (double*)M->vcycle_in(64), n);
    IMPES_LAUNCH_CHECK();
  };
  // z = M v (input already staged by normalize) and w = D^-1 A z
  auto precondition_apply = [&](const double* vj, double* zj) {
    if (!M) {
      launch_copy_f64(zj, vj, n, s);
      apply(A, zj, w_.p, s);
      return;
    }
    M->vcycle_only(precision, s);
    if (A.stencil) {
      if (precision == 32)
        launch_st7_rowscaled_spmv_z_f64(A.st, A.dinv, M->out_scale(), (const float*)M->vcycle_out(32), zj, w_.p, s);
      else
        launch_st7_rowscaled_spmv_z_f64(A.st, A.dinv, M->out_scale(), (const double*)M->vcycle_out(64), zj, w_.p,
                                        s);
      return;
    }
    if (precision == 32) launch_scale_f32_to_f64(zj, M->out_scale(), (const float*)M->vcycle_out(32), n, s);
    else launch_scale_f64(zj, M->out_scale(), (const double*)M->vcycle_out(64), n, s);
    apply(A, zj, w_.p, s);
  };
  residual(A, x, b, V, s);
  launch_dot_f64(V, V, n, sc + 4, work_.p, s);
  double beta = std::sqrt(read(sc + 4));
  st.r0_rel = beta / bnorm;
  int its = 0;
  const bool check = etol_ > 0.0 && M != nullptr;
  // residual target of the Krylov cycles: tol, or lower when the error check
  // predicts that tol leaves too large a pressure error
  double target = tol;
  const double ratio = std::max(err_ratio_[0], err_ratio_[1]);
  if (check && ratio > 0.0) target = std::min(tol, kErrTarget * etol_ / ratio * bnorm);
  bool ok = false;
  for (;;) {
    while (beta > target && its < maxit && std::isfinite(beta)) {
      k_init_cycle<<<1, 1, 0, s>>>(L, sm);
      IMPES_LAUNCH_CHECK();
      normalize(V, V, sc + 0);
      int k = 0;
      const int its0 = its;
      // the estimate of iteration jj lands in pinned slot host_[2 + (jj & 1)]; its
      // event is waited on later
      auto post = [&](int jj) { IMPES_HIP_CHECK(hipEventRecord(est_ev_[jj & 1], s)); };
      auto read_est = [&](int jj) {
        IMPES_HIP_CHECK(hipEventSynchronize(est_ev_[jj & 1]));
        return ((volatile double*)host_)[2 + (jj & 1)];
      };
      auto stops = [&](int jj, double est) { return !(est > target) || its0 + jj + 1 >= maxit; };
      // last two residual estimates read (the cycle's initial residual first)
      double e1 = beta, e2 = -1.0;
      int pending = -1;  // iteration whose estimate is still to be read
      for (int j = 0; j < m; ++j) {
        double* vj = V + (long long)j * n;
        double* zj = Z + (long long)j * n;
        double* est = host_dev_ + 2 + (j & 1);
        precondition_apply(vj, zj);
        // CGS2: h = V^T w; then w -= V h with h2 = V^T w in one sweep; then w -= V h2
        // with ||w||^2 and the rotations
        const int kj = j + 1;
        if (kj <= kMaxFusedGs) {
          launch_gs_pass<0>(kj, w_.p, V, n, nullptr, work_.p, s);
          k_fold_cols<<<kj, kGsThreads, 0, s>>>(work_.p, kMultiDotBlocks, sm + L.h());
          IMPES_LAUNCH_CHECK();
          launch_gs_pass<1>(kj, w_.p, V, n, sm + L.h(), work_.p, s);
          k_fold_cols<<<kj, kGsThreads, 0, s>>>(work_.p, kMultiDotBlocks, sm + L.h2());
          IMPES_LAUNCH_CHECK();
          launch_gs_pass<2>(kj, w_.p, V, n, sm + L.h2(), work_.p, s);
        } else {
          launch_multidot_f64(V, n, kj, w_.p, n, sm + L.h(), work_.p, s);
          launch_gs_update_f64(w_.p, V, n, kj, sm + L.h(), n, s);
          launch_multidot_f64(V, n, kj, w_.p, n, sm + L.h2(), work_.p, s);
          launch_gs_update_f64(w_.p, V, n, kj, sm + L.h2(), n, s);
          launch_dot_partial_f64(w_.p, w_.p, n, work_.p, kMultiDotBlocks, s);
        }
        k_arnoldi<<<1, kGsThreads, 0, s>>>(L, j, sm, work_.p, kMultiDotBlocks, est);
        IMPES_LAUNCH_CHECK();
        normalize(V + (long long)(j + 1) * n, w_.p, sc + 3);
        post(j);
        if (pending >= 0) {  // iteration j - 1 was not tested before j was enqueued
          const double e = read_est(pending);
          e2 = e1;
          e1 = e;
          pending = -1;
          if (stops(j - 1, e)) {  // iteration j was speculative
            k = j;
            break;
          }
        }
        k = j + 1;
        // Test iteration j only after enqueuing j + 1 while it cannot end the cycle
        // by prediction: the last estimate times the last contraction stays above
        // kDeferMargin x tol. The GPU then does not wait for the host between the
        // two iterations; a wrong prediction costs one ignored iteration.
        bool defer = !no_defer_ && j < m - 1 && its0 + j + 2 < maxit;
        if (defer && !lookahead_) defer = e2 > 0.0 && e1 * std::min(1.0, e1 / e2) > kDeferMargin * target;
        if (defer) {
          pending = j;
          continue;
        }
        const double e = read_est(j);
        e2 = e1;
        e1 = e;
        if (stops(j, e)) break;
      }
      its = its0 + k;
      k_backsolve<<<1, 1, 0, s>>>(L, k, sm);
      IMPES_LAUNCH_CHECK();
      launch_lincomb_add_f64(x, Z, n, k, sm + L.y(), n, s);
      residual(A, x, b, V, s);
      launch_dot_f64(V, V, n, sc + 4, work_.p, s);
      beta = std::sqrt(read(sc + 4));
      ++st.restarts;
    }
    if (!check || !(beta <= tol)) {
      ok = beta <= tol;
      break;
    }
    // error check: V holds the scaled residual D^-1 (b - A x); z = M V lands in
    // Z's first column (w receives D^-1 A z, unused)
    launch_fill_f64(sc + 6, 1.0, 1, s);
    normalize(w_.p, V, sc + 6);
    precondition_apply(w_.p, Z);
    launch_amax_f64(Z, n, sc + 6, work_.p, s);
    launch_axpy_f64(x, 1.0, Z, n, s);
    launch_amax_f64(x, n, sc + 7, work_.p, s);
    residual(A, x, b, V, s);
    launch_dot_f64(V, V, n, sc + 4, work_.p, s);
    IMPES_HIP_CHECK(hipMemcpyAsync(host_ + 4, sc + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
    const double zmax = host_[6], xmax = host_[7];
    const double est = (zmax > 0.0) ? zmax / std::max(xmax, 1e-300) : 0.0;
    err_ratio_[1] = err_ratio_[0];
    err_ratio_[0] = est / std::max(beta / bnorm, 1e-300);
    beta = std::sqrt(host_[4]);
    ++st.checks;
    st.err_estimate = est;
    ok = est <= kErrAccept * etol_ && beta <= tol;
    if (ok || its >= maxit || !std::isfinite(beta)) break;
    target = std::min(target, beta * kErrTarget * etol_ / std::max(est, 1e-300));
  }
  st.iterations = its;
  st.rel_residual = beta / bnorm;
  st.converged = ok ? 1 : 0;
}

}  // namespace impes
