// Dense inverse of a large coarsest AMG operator (classical_amg.py:1045-1069 solves
// the coarsest level densely) on the matrix cores: blocked Gauss-Jordan without
// pivoting, whose rank-64 updates run as 64 x 64 x 64 FP64 tiles of
// v_mfma_f64_16x16x4_f64. The operator is symmetric positive definite (a Galerkin
// product of the SPD pressure matrix, with DOF 0 optionally pinned), so every
// diagonal block met during the elimination is a Schur complement of an SPD
// matrix and itself SPD: no pivoting is needed.
//
// The matrix (n x n, row-major) is padded to n_pad = 64 nb with an identity block.
// For each block column k (one step of Gauss-Jordan on 64 columns at once), two
// launches:
//   P = A_kk^-1                          one workgroup, the block in registers (4 x 4 pivot blocks)
//   from A into A':                      one workgroup per 64 x 64 block, MFMA
// where the first P comes from its own launch and each later one from the
// workgroup of the previous step that produces that pivot block (so one launch
// per block column after the first)
//     A'_kk = P;  A'_kj = P A_kj;  A'_ik = -A_ik P;  A'_ij = A_ij - A_ik (P A_kj)
// which is the scalar elimination (a_kj /= a_kk; a_ij -= a_ik a_kj; a_ik = -a_ik /
// a_kk; a_kk = 1 / a_kk) on 64 x 64 blocks. Writing into a second matrix lets
// every block read the old row and column panels, so each block forms its own
// P A_kj instead of waiting on a row-panel launch. After the last block column
// the inverse is in M (the elimination starts in W when nb is odd). Each visit of
// the V-cycle then applies the inverse as one dense matrix-vector product (a
// wavefront per row, FP64 accumulation).
#include <utility>

#include "amg_kernels.h"

namespace impes {

namespace {

constexpr int kB = 64;        // block size of the elimination
constexpr int kLd = kB + 1;   // LDS row stride (doubles): rows fall on distinct banks
constexpr int kGemmThreads = 256;
typedef double v4d __attribute__((ext_vector_type(4)));

// C = A B for 64 x 64 tiles in LDS (row stride kLd). Wave w owns rows 16w..16w+15
// as four 16 x 16 output tiles t (columns 16t..16t+15). v_mfma_f64_16x16x4_f64 lane
// layout (measured on gfx950): A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
// D[i][j] at lane j + 16 (i % 4), element i / 4.
__device__ __forceinline__ void tile_mma(const double* __restrict__ As, const double* __restrict__ Bs, v4d acc[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kk = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k0 = 0; k0 < kB; k0 += 4) {
    const double a = As[(16 * w + i) * kLd + k0 + kk];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double b = Bs[(k0 + kk) * kLd + 16 * t + i];
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
  }
}

// (row, column) within the 64 x 64 tile of element e of output tile t
__device__ __forceinline__ void tile_pos(int t, int e, int& r, int& c) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  r = 16 * w + 4 * e + (lane >> 4);
  c = 16 * t + (lane & 15);
}

// global 64 x 64 block (row-major, leading dimension ld) -> LDS tile
__device__ __forceinline__ void load_tile(const double* __restrict__ g, long long ld, double* __restrict__ s) {
  for (int q = threadIdx.x; q < kB * kB; q += blockDim.x) {
    const int r = q / kB, c = q % kB;
    s[r * kLd + c] = g[r * ld + c];
  }
}

// Dense n_pad x n_pad from CSR: one workgroup per row; the row is staged in LDS,
// its entries loaded by all threads a block at a time and added in entry order
// by one thread (repeated columns add up as in a serial scatter), then written
// out. Padding rows are identity rows; a pinned DOF 0 keeps only a_00 = 1 in its
// row and column.
constexpr int kFillThreads = 256;
__global__ __launch_bounds__(kFillThreads) void k_fill_padded(int n, int n_pad, const int* __restrict__ rp,
                                                              const int* __restrict__ ci, const double* __restrict__ v,
                                                              const int* __restrict__ pin_flag, double* __restrict__ M) {
  extern __shared__ double row[];
  __shared__ int sc[kFillThreads];
  __shared__ double sv[kFillThreads];
  const int r = blockIdx.x;
  const int pin = *pin_flag;
  for (int c = threadIdx.x; c < n_pad; c += blockDim.x) row[c] = 0.0;
  if (r >= n || (pin && r == 0)) {
    __syncthreads();
    if (threadIdx.x == 0) row[r] = 1.0;
  } else {
    const int e0 = rp[r], e1 = rp[r + 1];
    for (int b = e0; b < e1; b += kFillThreads) {
      const int e = b + threadIdx.x;
      if (e < e1) {
        sc[threadIdx.x] = ci[e];
        sv[threadIdx.x] = v[e];
      }
      __syncthreads();  // also orders the zeroing above before the first scatter
      if (threadIdx.x == 0) {
        const int m = min(kFillThreads, e1 - b);
        for (int q = 0; q < m; ++q) {
          const int c = sc[q];
          if (pin && c == 0) continue;
          row[c] += sv[q];
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n_pad; c += blockDim.x) M[(long long)r * n_pad + c] = row[c];
}

// P = A_kk^-1 by unpivoted Gauss-Jordan on 4 x 4 pivot blocks, the 64 x 64 block
// in registers: thread t holds column j = t % 64 at rows i = t / 64 + 4 m
// (m < 16). Step q (pivot rows and columns 4q..4q+3) exchanges only the old
// pivot columns and rows through LDS (double-buffered: one barrier per step);
// every thread inverts the 4 x 4 pivot block B itself (scalar Gauss-Jordan, in
// registers) and updates its 16 entries with the block elimination's formulas:
//   pivot rows: B^-1 A_Pj (B^-1 inside the pivot block); pivot columns outside
//   the pivot rows: -A_iP B^-1; else A_ij - A_iP (B^-1 A_Pj).
// The reads of a step issue before the branch-free update that uses them. 16
// barrier steps instead of 64: 88 -> 28 us per 64 x 64 block. A non-positive
// pivot of a 4 x 4 inverse (not SPD) sets *bad.
// The inversion above on a block held as a[16] per thread (256 threads, thread t:
// column t % 64, rows t / 64 + 4 m); colb and rowb are the exchange buffers.
// Returns nonzero in a thread that met a non-positive pivot.
__device__ __forceinline__ int invert_block(double (&a)[16], double (&colb)[2][kB][4], double (&rowb)[2][4][kB]) {
  const int j = threadIdx.x & (kB - 1), i0 = threadIdx.x >> 6;
  int badp = 0;
  for (int q = 0; q < kB / 4; ++q) {
    const int buf = q & 1;
    const int p0 = 4 * q;
    const bool jp = (j >> 2) == q;  // this column is a pivot column
    if (jp) {
#pragma unroll
      for (int m = 0; m < 16; ++m) colb[buf][i0 + 4 * m][j - p0] = a[m];
    }
    {
      double v = 0.0;  // this thread's entry of pivot row p0 + i0 (its m = q)
#pragma unroll
      for (int m = 0; m < 16; ++m) v = (m == q) ? a[m] : v;
      rowb[buf][i0][j] = v;
    }
    __syncthreads();
    double B[4][4], rc[4], cc[16][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int c = 0; c < 4; ++c) B[r][c] = rowb[buf][r][p0 + c];
      rc[r] = rowb[buf][r][j];
    }
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) cc[m][t] = colb[buf][i0 + 4 * m][t];
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) {  // B <- B^-1
      const double piv = B[pp][pp];
      badp |= !(piv > 0.0);
      const double inv = 1.0 / piv;
#pragma unroll
      for (int c = 0; c < 4; ++c) B[pp][c] = (c == pp) ? inv : B[pp][c] * inv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r == pp) continue;
        const double f = B[r][pp];
#pragma unroll
        for (int c = 0; c < 4; ++c) B[r][c] = (c == pp) ? -f * inv : B[r][c] - f * B[pp][c];
      }
    }
    double R[4], Bj[4];  // the new pivot rows in this column; column j - p0 of B^-1
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = 0.0, bj = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        v += B[r][c] * rc[c];
        bj = (j - p0 == c) ? B[r][c] : bj;
      }
      Bj[r] = bj;
      R[r] = jp ? bj : v;
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      double sR = 0.0, sB = 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        sR += cc[m][t] * R[t];
        sB += cc[m][t] * Bj[t];
      }
      const double outside = jp ? -sB : a[m] - sR;
      a[m] = (m == q) ? R[i0] : outside;
    }
  }
  return badp;
}

__global__ __launch_bounds__(256) void k_gj_diag(int k, int n_pad, const double* __restrict__ M,
                                                 double* __restrict__ P, int* __restrict__ bad) {
  __shared__ double colb[2][kB][4], rowb[2][4][kB];
  const int j = threadIdx.x & (kB - 1), i0 = threadIdx.x >> 6;
  const double* src = M + (long long)k * kB * n_pad + (long long)k * kB;
  double a[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) a[m] = src[(long long)(i0 + 4 * m) * n_pad + j];
  const int badp = invert_block(a, colb, rowb);
  if (threadIdx.x == 0 && badp) *bad = 1;
#pragma unroll
  for (int m = 0; m < 16; ++m) P[(i0 + 4 * m) * kB + j] = a[m];
}

// One elimination step from A into A' (see the header), one workgroup per block
// (blockIdx.x = j, blockIdx.y = i). Every block of column j != k forms P A_kj
// itself with the same MFMA sequence, so the row panel it subtracts equals the one
// its own row-k block stores.
__global__ __launch_bounds__(kGemmThreads) void k_gj_step(int k, int n_pad, const double* __restrict__ A,
                                                          double* __restrict__ An, const double* __restrict__ P,
                                                          double* __restrict__ Pn, int* __restrict__ bad) {
  __shared__ double As[kB * kLd], Bs[kB * kLd];
  __shared__ double colb[2][kB][4], rowb[2][4][kB];
  const int j = blockIdx.x, i = blockIdx.y;
  const double* blk = A + (long long)i * kB * n_pad + (long long)j * kB;
  double* out = An + (long long)i * kB * n_pad + (long long)j * kB;
  if (i == k && j == k) {
    for (int q = threadIdx.x; q < kB * kB; q += blockDim.x) out[(long long)(q / kB) * n_pad + q % kB] = P[q];
    return;
  }
  v4d acc[4];
  if (j == k) {  // A'_ik = -A_ik P
    load_tile(blk, n_pad, As);
    load_tile(P, kB, Bs);
    __syncthreads();
    tile_mma(As, Bs, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r, c;
        tile_pos(t, e, r, c);
        out[(long long)r * n_pad + c] = -acc[t][e];
      }
    return;
  }
  load_tile(P, kB, As);
  load_tile(A + (long long)k * kB * n_pad + (long long)j * kB, n_pad, Bs);
  __syncthreads();
  tile_mma(As, Bs, acc);  // P A_kj
  if (i == k) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r, c;
        tile_pos(t, e, r, c);
        out[(long long)r * n_pad + c] = acc[t][e];
      }
    return;
  }
  __syncthreads();  // every wave is done reading As and Bs
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int r, c;
      tile_pos(t, e, r, c);
      Bs[r * kLd + c] = acc[t][e];
    }
  load_tile(A + (long long)i * kB * n_pad + (long long)k * kB, n_pad, As);
  __syncthreads();
  tile_mma(As, Bs, acc);  // A_ik (P A_kj)
  const bool next_pivot = (i == k + 1 && j == k + 1 && Pn != nullptr);
  if (next_pivot) __syncthreads();  // every wave is done reading As (block-uniform branch)
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int r, c;
      tile_pos(t, e, r, c);
      const double v = blk[(long long)r * n_pad + c] - acc[t][e];
      out[(long long)r * n_pad + c] = v;
      if (next_pivot) As[r * kLd + c] = v;
    }
  if (!next_pivot) return;
  // this block is the next step's pivot block: invert it here (as k_gj_diag would
  // from the stored block) and hand the inverse to step k + 1 in Pn
  __syncthreads();
  const int jj = threadIdx.x & (kB - 1), i0 = threadIdx.x >> 6;
  double a[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) a[m] = As[(i0 + 4 * m) * kLd + jj];
  const int badp = invert_block(a, colb, rowb);
  if (threadIdx.x == 0 && badp) *bad = 1;
#pragma unroll
  for (int m = 0; m < 16; ++m) Pn[(i0 + 4 * m) * kB + jj] = a[m];
}

// x = A^-1 b (b_0 taken as 0 when DOF 0 is pinned): a wavefront per row, four
// independent partial sums per lane (four loads in flight), folded at the end.
template <class T>
__global__ __launch_bounds__(256) void k_dense_mv(int n, int n_pad, const int* __restrict__ pin_flag,
                                                  const double* __restrict__ M, const T* __restrict__ b,
                                                  T* __restrict__ x) {
  const int pin = *pin_flag;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nwaves = (gridDim.x * blockDim.x) / kWave;
  for (int r = wid; r < n; r += nwaves) {
    const double* row = M + (long long)r * n_pad;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = (pin ? 1 : 0) + lane;
    for (; c + 3 * kWave < n; c += 4 * kWave) {
      s0 += row[c] * (double)b[c];
      s1 += row[c + kWave] * (double)b[c + kWave];
      s2 += row[c + 2 * kWave] * (double)b[c + 2 * kWave];
      s3 += row[c + 3 * kWave] * (double)b[c + 3 * kWave];
    }
    for (; c < n; c += kWave) s0 += row[c] * (double)b[c];
    const double s = wave_sum((s0 + s1) + (s2 + s3));
    if (lane == 0) x[r] = (T)s;
  }
}

}  // namespace

int dense_pad(int n) { return div_up(n, kB) * kB; }

void launch_dense_inverse_mfma(int n, const int* rp, const int* ci, const double* v, const int* pin, double* M,
                               double* W, double* P, int* bad, hipStream_t s) {
  IMPES_REQUIRE(n >= 1 && n <= kDenseMfmaMax, "MFMA dense inverse supports 1 <= n <= 4096");
  const int n_pad = dense_pad(n), nb = n_pad / kB;
  // each step swaps the two matrices: start where nb steps end in M
  double* cur = (nb % 2 == 0) ? M : W;
  double* nxt = (nb % 2 == 0) ? W : M;
  k_fill_padded<<<n_pad, kFillThreads, n_pad * sizeof(double), s>>>(n, n_pad, rp, ci, v, pin, cur);
  IMPES_LAUNCH_CHECK();
  // P holds two 64 x 64 pivot inverses: step k reads P_k and its block (k+1, k+1)
  // workgroup writes P_{k+1} into the other half
  // (a separate pivot-inverse launch per block column took 7.72-7.73 against
  // 7.67-7.70 ms per step at 1M cells)
  k_gj_diag<<<1, 256, 0, s>>>(0, n_pad, cur, P, bad);
  IMPES_LAUNCH_CHECK();
  for (int k = 0; k < nb; ++k) {
    double* Pk = P + (k & 1) * kB * kB;
    double* Pn = (k + 1 < nb) ? P + ((k + 1) & 1) * kB * kB : nullptr;
    k_gj_step<<<dim3(nb, nb), kGemmThreads, 0, s>>>(k, n_pad, cur, nxt, Pk, Pn, bad);
    IMPES_LAUNCH_CHECK();
    std::swap(cur, nxt);
  }
}

void launch_dense_mv_f64(int n, const int* pin, const double* M, const double* b, double* x, hipStream_t s) {
  k_dense_mv<double><<<div_up(n, 4), 256, 0, s>>>(n, dense_pad(n), pin, M, b, x);
  IMPES_LAUNCH_CHECK();
}

void launch_dense_mv_f32(int n, const int* pin, const double* M, const float* b, float* x, hipStream_t s) {
  k_dense_mv<float><<<div_up(n, 4), 256, 0, s>>>(n, dense_pad(n), pin, M, b, x);
  IMPES_LAUNCH_CHECK();
}

}  // namespace impes
