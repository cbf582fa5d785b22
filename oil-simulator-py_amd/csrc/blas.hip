// Level-1 vector kernels for the FP64 Krylov solver (FGMRES with classical
// Gram-Schmidt + reorthogonalisation) and the FP32/FP64 precision boundary.
//
// Reductions are two-stage and deterministic: a fixed grid of kRedBlocks
// blocks writes partial sums in block order, one block folds them. No atomics,
// so repeated runs give bit-identical results.
#include "kernels.h"

namespace impes {

namespace {
constexpr int kBlock = 256;

template <class T>
__global__ void k_fill(T* __restrict__ x, T v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

template <class TO, class TI>
__global__ void k_copy(TO* __restrict__ y, const TI* __restrict__ x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}

template <class TO, class TI>
__global__ void k_scale(TO* __restrict__ y, const double* __restrict__ a, const TI* __restrict__ x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = (TO)(a[i] * (double)x[i]);
}

__global__ void k_axpy(double* __restrict__ y, double alpha, const double* __restrict__ x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

__global__ void k_dot_partial(const double* __restrict__ x, const double* __restrict__ y, long long n,
                              double* __restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_amax_partial(const double* __restrict__ x, long long n, double* __restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s = fmax(s, fabs(x[i]));
  s = block_max(s, lds, 0.0);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// fold `nb` partials per column for k columns: out[c] = sum_b part[c*nb + b]
template <bool MAX>
__global__ void k_fold(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double lds[kBlock / kWave];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    double v = part[(long long)c * nb + b];
    s = MAX ? fmax(s, v) : s + v;
  }
  s = MAX ? block_max(s, lds, 0.0) : block_sum(s, lds);
  if (threadIdx.x == 0) out[c] = s;
}

// Partial dots of w against k basis vectors. Each thread keeps its w element in
// a register and streams the k basis entries (coalesced along the row).
__global__ void k_multidot_partial(const double* __restrict__ V, long long ldv, int k,
                                   const double* __restrict__ w, long long n, double* __restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  // Process basis vectors in chunks of 8 so the per-thread accumulators stay in registers.
  for (int c0 = 0; c0 < k; c0 += 8) {
    const int cn = min(8, k - c0);
    double acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      const double wi = w[i];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (t < cn) acc[t] += V[(long long)(c0 + t) * ldv + i] * wi;
    }
    for (int t = 0; t < cn; ++t) {
      double s = block_sum(acc[t], lds);
      if (threadIdx.x == 0) part[(long long)(c0 + t) * gridDim.x + blockIdx.x] = s;
    }
  }
}

__global__ void k_gs_update(double* __restrict__ w, const double* __restrict__ V, long long ldv, int k,
                            const double* __restrict__ h, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double s = w[i];
    for (int c = 0; c < k; ++c) s -= V[(long long)c * ldv + i] * h[c];
    w[i] = s;
  }
}

__global__ void k_lincomb_add(double* __restrict__ x, const double* __restrict__ Z, long long ldz, int k,
                              const double* __restrict__ y, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double s = x[i];
    for (int c = 0; c < k; ++c) s += Z[(long long)c * ldz + i] * y[c];
    x[i] = s;
  }
}

__global__ void k_div(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ a,
                      long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i] / a[i];
}

__global__ void k_xpby(double* __restrict__ y, const double* __restrict__ x, double beta, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i] + beta * y[i];
}

// v may alias w (in-place normalisation)
__global__ void k_scale_by_inv(double* v, const double* w, const double* __restrict__ nrm,
                               long long n) {
  const double inv = 1.0 / nrm[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    v[i] = w[i] * inv;
}

}  // namespace

void launch_fill_f64(double* x, double v, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<double><<<grid_for(n, kBlock), kBlock, 0, s>>>(x, v, n);
  IMPES_LAUNCH_CHECK();
}
void launch_fill_f32(float* x, float v, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<float><<<grid_for(n, kBlock), kBlock, 0, s>>>(x, v, n);
  IMPES_LAUNCH_CHECK();
}
void launch_fill_i32(int* x, int v, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<int><<<grid_for(n, kBlock), kBlock, 0, s>>>(x, v, n);
  IMPES_LAUNCH_CHECK();
}
void launch_copy_f64(double* y, const double* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<double, double><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_copy_f32(float* y, const float* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<float, float><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_f32_to_f64(double* y, const float* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<double, float><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, n);
  IMPES_LAUNCH_CHECK();
}

void launch_f64_to_f32(float* y, const double* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<float, double><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_scale_f64_to_f32(float* y, const double* a, const double* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_scale<float, double><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, a, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_scale_f32_to_f64(double* y, const double* a, const float* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_scale<double, float><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, a, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_scale_f64(double* y, const double* a, const double* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_scale<double, double><<<grid_for(n, kBlock), kBlock, 0, s>>>(y, a, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_axpy_f64(double* y, double alpha, const double* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_axpy<<<grid_for(n, kBlock), kBlock, 0, s>>>(y, alpha, x, n);
  IMPES_LAUNCH_CHECK();
}
void launch_div_f64(double* y, const double* x, const double* a, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_div<<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, a, n);
  IMPES_LAUNCH_CHECK();
}
void launch_xpby_f64(double* y, const double* x, double beta, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_xpby<<<grid_for(n, kBlock), kBlock, 0, s>>>(y, x, beta, n);
  IMPES_LAUNCH_CHECK();
}
void launch_dot_partial_f64(const double* x, const double* y, long long n, double* part, int nb, hipStream_t s) {
  k_dot_partial<<<nb, kBlock, 0, s>>>(x, y, n, part);
  IMPES_LAUNCH_CHECK();
}
void launch_dot_f64(const double* x, const double* y, long long n, double* out, double* work, hipStream_t s) {
  k_dot_partial<<<kRedBlocks, kBlock, 0, s>>>(x, y, n, work);
  IMPES_LAUNCH_CHECK();
  k_fold<false><<<1, kBlock, 0, s>>>(work, kRedBlocks, out);
  IMPES_LAUNCH_CHECK();
}
void launch_amax_f64(const double* x, long long n, double* out, double* work, hipStream_t s) {
  k_amax_partial<<<kRedBlocks, kBlock, 0, s>>>(x, n, work);
  IMPES_LAUNCH_CHECK();
  k_fold<true><<<1, kBlock, 0, s>>>(work, kRedBlocks, out);
  IMPES_LAUNCH_CHECK();
}
void launch_multidot_f64(const double* V, long long ldv, int k, const double* w, long long n, double* h,
                         double* work, hipStream_t s) {
  IMPES_REQUIRE(k >= 1 && k <= kMaxMultiDot, "multidot: k out of range");
  // more waves in flight than the scalar reductions: this one streams k vectors
  k_multidot_partial<<<kMultiDotBlocks, kBlock, 0, s>>>(V, ldv, k, w, n, work);
  IMPES_LAUNCH_CHECK();
  k_fold<false><<<k, kBlock, 0, s>>>(work, kMultiDotBlocks, h);
  IMPES_LAUNCH_CHECK();
}
void launch_gs_update_f64(double* w, const double* V, long long ldv, int k, const double* h, long long n,
                          hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  k_gs_update<<<grid_for(n, kBlock), kBlock, 0, s>>>(w, V, ldv, k, h, n);
  IMPES_LAUNCH_CHECK();
}
void launch_lincomb_add_f64(double* x, const double* Z, long long ldz, int k, const double* y, long long n,
                            hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  k_lincomb_add<<<grid_for(n, kBlock), kBlock, 0, s>>>(x, Z, ldz, k, y, n);
  IMPES_LAUNCH_CHECK();
}
void launch_scale_by_inv_f64(double* v_out, const double* w, const double* nrm, long long n, hipStream_t s) {
  if (n <= 0) return;
  k_scale_by_inv<<<grid_for(n, kBlock), kBlock, 0, s>>>(v_out, w, nrm, n);
  IMPES_LAUNCH_CHECK();
}

}  // namespace impes
