// FP64 restarted FGMRES(m) with CGS2 orthogonalisation. The Hessenberg matrix,
// Givens rotations and the residual estimate live on the device; the host reads
// one scalar per Arnoldi step to decide termination.
#include "fgmres.h"

#include <algorithm>
#include <cstdlib>

#include <cmath>

namespace impes {

namespace {

constexpr int kArnoldiBlock = 256;

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

// Column j of H from the two Gram-Schmidt passes, rotations, residual estimate.
// The block first folds the nb partial sums of ||w||^2 left by the second pass.
__global__ void k_arnoldi(Layout L, int j, double* __restrict__ sm, const double* __restrict__ part, int nb) {
  __shared__ double lds[kArnoldiBlock / kWave];
  double ww = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) ww += part[b];
  ww = block_sum(ww, lds);
  if (threadIdx.x != 0) return;
  double* sc = sm + L.sc();
  sc[1] = ww;
  const double* h = sm + L.h();
  const double* h2 = sm + L.h2();
  double* cs = sm + L.cs();
  double* sn = sm + L.sn();
  double* g = sm + L.g();
  for (int i = 0; i <= j; ++i) sm[L.H(i, j)] = h[i] + h2[i];
  const double hn = sqrt(fmax(sc[1], 0.0));
  sm[L.H(j + 1, j)] = hn;
  for (int i = 0; i < j; ++i) {
    const double a = sm[L.H(i, j)], b = sm[L.H(i + 1, j)];
    sm[L.H(i, j)] = cs[i] * a + sn[i] * b;
    sm[L.H(i + 1, j)] = -sn[i] * a + cs[i] * b;
  }
  const double a = sm[L.H(j, j)], b = sm[L.H(j + 1, j)];
  double c, s;
  if (b == 0.0) { c = 1.0; s = 0.0; }
  else if (fabs(b) > fabs(a)) { const double t = a / b; s = 1.0 / sqrt(1.0 + t * t); c = s * t; }
  else { const double t = b / a; c = 1.0 / sqrt(1.0 + t * t); s = c * t; }
  cs[j] = c;
  sn[j] = s;
  sm[L.H(j, j)] = c * a + s * b;
  sm[L.H(j + 1, j)] = 0.0;
  g[j + 1] = -s * g[j];
  g[j] = c * g[j];
  sc[2] = fabs(g[j + 1]);
  sc[3] = (hn > 0.0) ? hn : 1.0;
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

Fgmres::Fgmres() {
  const char* e = std::getenv("IMPES_FGMRES_LOOKAHEAD");
  lookahead_ = e && e[0] == '1';
}

Fgmres::~Fgmres() {
  if (host_) (void)hipHostFree(host_);
  for (auto& e : est_ev_)
    if (e) (void)hipEventDestroy(e);
}

void Fgmres::ensure(long long n, int m) {
  if (!host_) IMPES_HIP_CHECK(hipHostMalloc((void**)&host_, 8 * sizeof(double), hipHostMallocDefault));
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
  residual(A, x, b, V, s);
  launch_dot_f64(V, V, n, sc + 4, work_.p, s);
  double beta = std::sqrt(read(sc + 4));
  st.r0_rel = beta / bnorm;
  int its = 0;
  while (beta > tol && its < maxit && std::isfinite(beta)) {
    k_init_cycle<<<1, 1, 0, s>>>(L, sm);
    IMPES_LAUNCH_CHECK();
    launch_scale_by_inv_f64(V, V, sc + 0, n, s);
    int k = 0;
    const int its0 = its;
    // estimate of iteration jj: async copy into a pinned slot + event, waited on later
    auto post = [&](int jj) {
      IMPES_HIP_CHECK(hipMemcpyAsync(host_ + 2 + (jj & 1), sc + 2, sizeof(double), hipMemcpyDeviceToHost, s));
      IMPES_HIP_CHECK(hipEventRecord(est_ev_[jj & 1], s));
    };
    auto done = [&](int jj) {  // iteration jj (0-based) ends the cycle?
      IMPES_HIP_CHECK(hipEventSynchronize(est_ev_[jj & 1]));
      const double est = host_[2 + (jj & 1)];
      return !(est > tol) || its0 + jj + 1 >= maxit;
    };
    for (int j = 0; j < m; ++j) {
      double* vj = V + (long long)j * n;
      double* zj = Z + (long long)j * n;
      if (M) M->precondition(precision, vj, zj, s);
      else launch_copy_f64(zj, vj, n, s);
      apply(A, zj, w_.p, s);
      // CGS2: h = V^T w; w -= V h and h2 = V^T w in one sweep; w -= V h2 with the
      // partial sums of ||w||^2, folded by k_arnoldi
      launch_multidot_f64(V, n, j + 1, w_.p, n, sm + L.h(), work_.p, s);
      launch_gs_update_multidot_f64(w_.p, V, n, j + 1, sm + L.h(), n, sm + L.h2(), work_.p, s);
      launch_gs_update_norm_f64(w_.p, V, n, j + 1, sm + L.h2(), n, work_.p, s);
      k_arnoldi<<<1, kArnoldiBlock, 0, s>>>(L, j, sm, work_.p, kMultiDotBlocks);
      IMPES_LAUNCH_CHECK();
      launch_scale_by_inv_f64(V + (long long)(j + 1) * n, w_.p, sc + 3, n, s);
      post(j);
      if (!lookahead_) {
        k = j + 1;
        if (done(j)) break;
        continue;
      }
      if (j > 0 && done(j - 1)) {  // iteration j was speculative
        k = j;
        break;
      }
      k = j + 1;
      if (j == m - 1) {
        (void)done(j);  // the cycle ends here anyway; wait for the copy slot
      }
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
  st.iterations = its;
  st.rel_residual = beta / bnorm;
  st.converged = (beta <= tol) ? 1 : 0;
}

}  // namespace impes
