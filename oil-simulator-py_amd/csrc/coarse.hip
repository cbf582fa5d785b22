// Coarsest-level direct solve of the V-cycle (classical_amg.py:1045-1069):
// the level matrix is made dense, DOF 0 is optionally pinned (row/column 0 zeroed,
// a_00 = 1, b_0 = 0), and the system is solved exactly. The pin exists for singular
// (pure Neumann) matrices; on a nonsingular but nearly singular operator it removes
// the near-nullspace correction, so the pin decision is a device flag computed from
// the level's row sums (launch_pin_decide). The inverse is formed once per hierarchy
// update by Gauss-Jordan elimination with partial pivoting in FP64, so each V-cycle
// visit is a single dense matrix-vector product.
#include "amg_kernels.h"

namespace impes {

namespace {

// One workgroup, 8 lanes per row: pin = mode==1 || (mode==2 && no row has
// |row sum| > 1e-8 |a_ii|).
constexpr int kPinThreads = 1024;
constexpr int kPinLanes = 8;
__global__ __launch_bounds__(kPinThreads) void k_pin_decide(int n, const int* __restrict__ rp,
                                                            const int* __restrict__ ci, const double* __restrict__ v,
                                                            int mode, int* __restrict__ pin) {
  const int g = threadIdx.x % kPinLanes;
  int regular = 0;
  for (int base = 0; base < n; base += kPinThreads / kPinLanes) {
    const int r = base + (int)threadIdx.x / kPinLanes;
    double rs = 0.0, d = 0.0;
    if (r < n) {
      for (int e = rp[r] + g; e < rp[r + 1]; e += kPinLanes) {
        rs += v[e];
        if (ci[e] == r) d += v[e];
      }
    }
    rs = group_sum<kPinLanes>(rs);
    d = group_sum<kPinLanes>(d);
    if (r < n && g == 0 && fabs(rs) > 1e-8 * fabs(d)) regular = 1;
  }
  regular = __syncthreads_or(regular);
  if (threadIdx.x == 0) *pin = (mode == 1) ? 1 : (mode == 2 ? !regular : 0);
}

__global__ void k_pin_rhs(const int* __restrict__ pin, double* __restrict__ b) {
  if (*pin) b[0] = 0.0;
}
__global__ void k_pin_rhs32(const int* __restrict__ pin, float* __restrict__ b) {
  if (*pin) b[0] = 0.f;
}

// One workgroup per row r of [A_pinned | I]: the row's entries pass through LDS
// in chunks and each thread sums the entries of its (at most four) columns in
// entry order, so repeated column indices add up as in a serial scatter. Pinned
// DOF 0: row and column 0 zero but a_00 = 1.
constexpr int kDenseMax = 512;
constexpr int kFillThreads = 256;
constexpr int kFillPer = 2 * kDenseMax / kFillThreads;
__global__ __launch_bounds__(kFillThreads) void k_dense_fill(int n, const int* __restrict__ rp,
                                                             const int* __restrict__ ci, const double* __restrict__ v,
                                                             const int* __restrict__ pin_flag,
                                                             double* __restrict__ aug) {
  __shared__ int lc[kDenseMax];
  __shared__ double lv[kDenseMax];
  const int r = blockIdx.x;
  const long long w = 2LL * n;
  const int pin = *pin_flag;
  const int e0 = rp[r], m = rp[r + 1] - e0;
  double acc[kFillPer];
#pragma unroll
  for (int q = 0; q < kFillPer; ++q) acc[q] = 0.0;
  for (int t0 = 0; t0 < m; t0 += kDenseMax) {
    const int mc = (m - t0 < kDenseMax) ? m - t0 : kDenseMax;
    __syncthreads();  // the previous chunk is consumed
    for (int t = threadIdx.x; t < mc; t += kFillThreads) {
      lc[t] = ci[e0 + t0 + t];
      lv[t] = v[e0 + t0 + t];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kFillPer; ++q) {
      const int c = threadIdx.x + q * kFillThreads;
      if (c >= n || (pin && (r == 0 || c == 0))) continue;
      for (int t = 0; t < mc; ++t)
        if (lc[t] == c) acc[q] += lv[t];
    }
  }
#pragma unroll
  for (int q = 0; q < kFillPer; ++q) {
    const int c = threadIdx.x + q * kFillThreads;
    if (c >= 2 * n) continue;
    double a;
    if (c >= n) a = (c - n == r) ? 1.0 : 0.0;
    else if (pin && (r == 0 || c == 0)) a = (r == 0 && c == 0) ? 1.0 : 0.0;
    else a = acc[q];
    aug[r * w + c] = a;
  }
}

// Symmetric positive definite coarsest operators (every Galerkin level of a
// symmetric M-matrix) are inverted in place in LDS by the sweep operator on the
// packed lower triangle: for each pivot k, a_ij -= a_ik a_kj / a_kk (i, j != k),
// a_ik /= a_kk, a_kk = -1/a_kk; after all pivots the triangle holds -A^-1. Without
// pivoting this is stable for SPD matrices (it is Gauss-Jordan on the symmetric
// form). One workgroup, the whole triangle in LDS (n <= kSweepMax), two barriers
// per pivot. A non-symmetric matrix or a non-positive pivot leaves done = 0 and
// the pivoted Gauss-Jordan below runs instead.
constexpr int kSweepMax = 199;  // 199*200/2 doubles = 159,200 B of LDS
constexpr int kSweepPer = (kSweepMax * (kSweepMax + 1) / 2 + 1023) / 1024;

__global__ __launch_bounds__(1024) void k_sweep_inverse(int n, double* __restrict__ aug, int* __restrict__ done) {
  __shared__ double T[kSweepMax * (kSweepMax + 1) / 2];
  __shared__ double dg[kSweepMax];    // original diagonal
  __shared__ double colk[kSweepMax];  // column k of the current pivot, contiguous (conflict-free reads)
  __shared__ int bad;
  const long long w = 2LL * n;
  const int tid = threadIdx.x;
  if (tid == 0) *done = 0;
  if (n > kSweepMax) return;
  if (tid == 0) bad = 0;
  for (int i = tid; i < n; i += blockDim.x) dg[i] = aug[i * w + i];
  __syncthreads();
  // symmetric with a positive diagonal? |a_ij - a_ji| <= 1e-10 sqrt(a_ii a_jj): the
  // rounding asymmetry of a Galerkin product passes, a non-symmetric operator does not
  // (well rows can exceed interior rows by 15 orders, so the scale is per entry)
  int local_bad = 0;
  for (int i = tid; i < n; i += blockDim.x)
    if (!(dg[i] > 0.0)) local_bad = 1;
  for (int t = tid; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    if (j >= i) continue;
    const double d = fabs(aug[i * w + j] - aug[j * w + i]);
    if (!(d <= 1e-10 * sqrt(fabs(dg[i]) * fabs(dg[j])))) local_bad = 1;
  }
  if (local_bad) atomicOr(&bad, 1);
  __syncthreads();
  if (bad) return;  // uniform: every thread reads the same LDS flag
  // packed lower triangle: index i(i+1)/2 + j, j <= i; each thread owns fixed slots
  const int tot = n * (n + 1) / 2;
  short oi[kSweepPer], oj[kSweepPer];
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < kSweepPer; ++q) {
    const int p = tid + q * (int)blockDim.x;
    if (p >= tot) break;
    int i = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= p) ++i;
    while (i * (i + 1) / 2 > p) --i;
    const int j = p - i * (i + 1) / 2;
    oi[q] = (short)i;
    oj[q] = (short)j;
    cnt = q + 1;
    T[p] = aug[i * w + j];
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int i = tid; i < n; i += blockDim.x) colk[i] = (i >= k) ? T[i * (i + 1) / 2 + k] : T[k * (k + 1) / 2 + i];
    __syncthreads();
    const double d = colk[k];
    if (!(d > 1e-13 * dg[k])) {  // Schur pivot ~0 relative to a_kk: not (numerically) SPD
      if (tid == 0) bad = 1;
      break;
    }
    const double inv = 1.0 / d;
    // the update reads the staged copy of column k, so the column itself can be
    // scaled in the same pass; compile-time slot loops keep the slots in registers
#pragma unroll
    for (int q = 0; q < kSweepPer; ++q) {
      if (q >= cnt) break;
      const int i = oi[q], j = oj[q];
      const int p = i * (i + 1) / 2 + j;
      if (i == k && j == k) T[p] = -inv;
      else if (i == k || j == k) T[p] *= inv;
      else T[p] -= colk[i] * colk[j] * inv;
    }
    __syncthreads();
  }
  __syncthreads();
  if (bad) return;
  for (int t = tid; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    const int a = i > j ? i : j, b = i > j ? j : i;
    aug[i * w + n + j] = -T[a * (a + 1) / 2 + b];
  }
  if (tid == 0) *done = 1;
}

// The same sweep with the matrix in registers (n <= 32 R): the 1024 threads form a
// 32 x 32 grid and thread (ti, tj) owns rows ti + 32a and columns tj + 32b
// (a, b < R) of the full symmetric matrix, loaded from the lower triangle so both
// halves start (and stay) bit-equal. Per pivot only column k passes through LDS,
// double-buffered so one barrier per pivot suffices; the entry updates are the
// expressions of k_sweep_inverse in the same order, so the inverse is identical.
// The packed-LDS form above reads and writes every entry through LDS each pivot.
template <int R>
__global__ __launch_bounds__(1024) void k_sweep_regs(int n, double* __restrict__ aug, int* __restrict__ done) {
  constexpr int S = 32 * R;
  __shared__ double dg[S];
  __shared__ double colb[2][S];  // column k of pivot k in colb[k & 1]; zero beyond n
  __shared__ int bad;
  const long long w = 2LL * n;
  const int tid = threadIdx.x;
  if (tid == 0) *done = 0;
  if (n > S) return;
  if (tid == 0) bad = 0;
  for (int i = tid; i < S; i += blockDim.x) {
    dg[i] = i < n ? aug[i * w + i] : 1.0;
    colb[0][i] = 0.0;
    colb[1][i] = 0.0;
  }
  __syncthreads();
  int local_bad = 0;
  for (int i = tid; i < n; i += blockDim.x)
    if (!(dg[i] > 0.0)) local_bad = 1;
  for (int t = tid; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t % n;
    if (j >= i) continue;
    const double d = fabs(aug[i * w + j] - aug[j * w + i]);
    if (!(d <= 1e-10 * sqrt(fabs(dg[i]) * fabs(dg[j])))) local_bad = 1;
  }
  if (local_bad) atomicOr(&bad, 1);
  __syncthreads();
  if (bad) return;  // uniform
  const int ti = tid / 32, tj = tid % 32;
  double T[R][R];
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ti + 32 * a, j = tj + 32 * b;
      const int hi = i > j ? i : j, lo = i > j ? j : i;
      T[a][b] = (i < n && j < n) ? aug[hi * w + lo] : 0.0;
    }
  for (int k = 0; k < n; ++k) {
    double* col = colb[k & 1];
    if (tj == (k & 31)) {
      const int bk = k >> 5;
#pragma unroll
      for (int a = 0; a < R; ++a) {
        const int i = ti + 32 * a;
#pragma unroll
        for (int b = 0; b < R; ++b)
          if (b == bk && i < n) col[i] = T[a][b];
      }
    }
    __syncthreads();
    const double d = col[k];
    if (!(d > 1e-13 * dg[k])) {  // uniform: every thread reads the same LDS value
      if (tid == 0) bad = 1;
      break;
    }
    const double inv = 1.0 / d;
    double ci[R], cj[R];
#pragma unroll
    for (int a = 0; a < R; ++a) ci[a] = col[ti + 32 * a];
#pragma unroll
    for (int b = 0; b < R; ++b) cj[b] = col[tj + 32 * b];
    if constexpr (R >= 5) {
      // at 5 x 5 slots the split form below no longer fits 128 VGPRs
#pragma unroll
      for (int a = 0; a < R; ++a) {
        const int i = ti + 32 * a;
#pragma unroll
        for (int b = 0; b < R; ++b) {
          const int j = tj + 32 * b;
          if (i == k && j == k) T[a][b] = -inv;
          else if (i == k || j == k) T[a][b] *= inv;
          else T[a][b] -= ci[a] * cj[b] * inv;
        }
      }
    } else {
      // Every entry takes the off-pivot update, then the owners of row and column
      // k overwrite theirs with T_ik * inv = ci * inv and T_kj * inv = cj * inv (the
      // matrix stays bit-symmetric), and -inv on the pivot: the same values as the
      // per-entry form above. Branch-free for all but one half-wave row and one lane
      // per half-wave, so a pivot costs one product and one multiply-add per entry.
#pragma unroll
      for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) T[a][b] -= ci[a] * cj[b] * inv;
      const int kk = k & 31, bk = k >> 5;
      if (tj == kk) {
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b)
            if (b == bk) T[a][b] = (ti + 32 * a == k) ? -inv : ci[a] * inv;
      }
      if (ti == kk) {
#pragma unroll
        for (int a = 0; a < R; ++a)
          if (a == bk) {
#pragma unroll
            for (int b = 0; b < R; ++b)
              if (tj + 32 * b != k) T[a][b] = cj[b] * inv;
          }
      }
    }
  }
  __syncthreads();
  if (bad) return;
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ti + 32 * a, j = tj + 32 * b;
      if (i < n && j < n) aug[i * w + n + j] = -T[a][b];
    }
  if (tid == 0) *done = 1;
}

// Single workgroup Gauss-Jordan with partial pivoting on the n x 2n augmented
// matrix (n <= 500); skipped when the sweep already produced the inverse.
__global__ __launch_bounds__(1024) void k_gauss_jordan(int n, double* __restrict__ aug, const int* __restrict__ done) {
  if (*done) return;
  __shared__ double fcol[512];
  __shared__ double red_v[1024 / kWave];
  __shared__ int red_i[1024 / kWave];
  __shared__ int piv_row;
  const long long w = 2LL * n;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (int k = 0; k < n; ++k) {
    // pivot search over rows k..n-1 of column k
    double best = -1.0;
    int bi = k;
    for (int r = k + threadIdx.x; r < n; r += blockDim.x) {
      const double a = fabs(aug[(long long)r * w + k]);
      if (a > best) { best = a; bi = r; }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, kWave);
      const int oi = __shfl_xor(bi, off, kWave);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (lane == 0) { red_v[wid] = best; red_i[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = -1.0;
      int i = k;
      for (int q = 0; q < (int)(blockDim.x / kWave); ++q)
        if (red_v[q] > b || (red_v[q] == b && red_i[q] < i)) { b = red_v[q]; i = red_i[q]; }
      piv_row = i;
    }
    __syncthreads();
    const int p = piv_row;
    if (p != k) {
      for (long long c = threadIdx.x; c < w; c += blockDim.x) {
        const double t = aug[(long long)k * w + c];
        aug[(long long)k * w + c] = aug[(long long)p * w + c];
        aug[(long long)p * w + c] = t;
      }
    }
    __syncthreads();
    double piv = aug[(long long)k * w + k];
    if (fabs(piv) < 1e-300) piv = 1e-300;  // singular pivot: keep the elimination finite
    const double inv = 1.0 / piv;
    __syncthreads();
    for (long long c = k + threadIdx.x; c < w; c += blockDim.x) aug[(long long)k * w + c] *= inv;
    for (int r = threadIdx.x; r < n; r += blockDim.x) fcol[r] = (r == k) ? 0.0 : aug[(long long)r * w + k];
    __syncthreads();
    const long long span = w - k;
    for (long long t = threadIdx.x; t < (long long)n * span; t += blockDim.x) {
      const int r = (int)(t / span);
      const long long c = k + t % span;
      const double f = fcol[r];
      if (f != 0.0) aug[(long long)r * w + c] -= f * aug[(long long)k * w + c];
    }
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_dense_solve(int n, const int* __restrict__ pin_flag,
                                                     const double* __restrict__ aug, const T* __restrict__ b,
                                                     T* __restrict__ x) {
  const long long w = 2LL * n;
  const int pin = *pin_flag;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nwaves = (gridDim.x * blockDim.x) / kWave;
  for (int r = wid; r < n; r += nwaves) {
    double s = 0.0;
    for (int c = (pin ? 1 : 0) + lane; c < n; c += kWave) s += aug[(long long)r * w + n + c] * (double)b[c];
    s = wave_sum(s);
    if (lane == 0) x[r] = (T)s;
  }
}

}  // namespace

void launch_pin_decide(int n, const int* rp, const int* ci, const double* v, int mode, int* pin, hipStream_t s) {
  k_pin_decide<<<1, kPinThreads, 0, s>>>(n, rp, ci, v, mode, pin);
  IMPES_LAUNCH_CHECK();
}

void launch_pin_rhs_f64(const int* pin, double* b, hipStream_t s) {
  k_pin_rhs<<<1, 1, 0, s>>>(pin, b);
  IMPES_LAUNCH_CHECK();
}

void launch_pin_rhs_f32(const int* pin, float* b, hipStream_t s) {
  k_pin_rhs32<<<1, 1, 0, s>>>(pin, b);
  IMPES_LAUNCH_CHECK();
}

void launch_csr_to_dense(int n, const int* rp, const int* ci, const double* v, const int* pin, double* aug,
                         hipStream_t s) {
  IMPES_REQUIRE(n >= 1 && n <= 512, "dense coarsest solve supports 1 <= n <= 512");
  k_dense_fill<<<n, kFillThreads, 0, s>>>(n, rp, ci, v, pin, aug);
  IMPES_LAUNCH_CHECK();
}

void launch_gauss_jordan_inverse(int n, double* aug, int* done, hipStream_t s) {
  IMPES_REQUIRE(n >= 1 && n <= 512, "dense coarsest solve supports 1 <= n <= 512");
  // sweep in registers up to 32 x 5 rows (the tile must fit 128 VGPRs of a
  // 1024-thread block), in packed LDS up to kSweepMax, else Gauss-Jordan only
  if (n <= 64) k_sweep_regs<2><<<1, 1024, 0, s>>>(n, aug, done);
  else if (n <= 128) k_sweep_regs<4><<<1, 1024, 0, s>>>(n, aug, done);
  else if (n <= 160) k_sweep_regs<5><<<1, 1024, 0, s>>>(n, aug, done);
  else k_sweep_inverse<<<1, 1024, 0, s>>>(n, aug, done);
  IMPES_LAUNCH_CHECK();
  k_gauss_jordan<<<1, 1024, 0, s>>>(n, aug, done);
  IMPES_LAUNCH_CHECK();
}

void launch_dense_solve_f64(int n, const int* pin, const double* aug, const double* b, double* x, hipStream_t s) {
  k_dense_solve<double><<<div_up(n, 4), 256, 0, s>>>(n, pin, aug, b, x);
  IMPES_LAUNCH_CHECK();
}

void launch_dense_solve_f32(int n, const int* pin, const double* aug, const float* b, float* x, hipStream_t s) {
  k_dense_solve<float><<<div_up(n, 4), 256, 0, s>>>(n, pin, aug, b, x);
  IMPES_LAUNCH_CHECK();
}

}  // namespace impes
