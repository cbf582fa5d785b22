// Sparse operator kernels of the V-cycle and of the outer Krylov method.
//
// CSR: G lanes of a wavefront share one row (G = 1..64 chosen from the mean row
// length), lanes walk the row with stride G so a group's loads are contiguous,
// and the partial sums fold with cross-lane shuffles. Every variant is fused
// with the vector update it feeds (residual, L1-Jacobi sweep, prolongation
// add, row scaling) so each sweep streams the matrix once.
//
// Stencil7: the level-0 IMPES pressure operator is a symmetric 7-point stencil,
// so it is stored index-free as one diagonal and three face-coefficient arrays
// (4 values per cell instead of CSR's ~7 values + 7 indices + row pointers).
#include <cstdint>
#include <cstdlib>
#include <type_traits>

#include "kernels.h"
#include "ktimer.h"

namespace impes {

namespace {
constexpr int kBlock = 256;

// kRestrictJ0: y = A x (a restriction R r) and y2 = w * aux .* y (the first
// L1-Jacobi sweep of the coarse level from a zero guess, aux = its inv_relax)
// kYPlusB: y = b + A x (an add into another buffer; same arithmetic as kYAdd)
// kJacZ0 (stencil): two L1-Jacobi sweeps from a zero guess in one pass: the first
// sweep's x = w * aux .* b is formed at every stencil point instead of being read
// (the expression of k_jacobi0), then y = x + w * aux .* (b - A x) as kJac
enum Mode {
  kY = 0, kYAdd = 1, kRes = 2, kJac = 3, kRowScaled = 4, kRowScaledRes = 5, kRestrictJ0 = 6, kYPlusB = 7, kJacZ0 = 8
};

template <int G, class T, class TA, int MODE>
__global__ __launch_bounds__(kBlock) void k_csr(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const T* __restrict__ v, const T* __restrict__ x,
                                                 const T* __restrict__ b, const TA* __restrict__ aux, T w,
                                                 T* __restrict__ y, T* __restrict__ y2) {
  constexpr int RPB = kBlock / G;  // rows per block
  const int g = threadIdx.x % G;
  for (long long base = xcd_block() * RPB; base < n; base += (long long)gridDim.x * RPB) {
    const long long row = base + threadIdx.x / G;
    const bool active = row < n;
    // the row's own operands are loaded before the row walk, so their latency
    // overlaps the (row pointer -> column -> gather) chain instead of following it
    T ob = T(0), ox = T(0), oy = T(0), oa = T(0);
    if (active && g == 0) {
      if (MODE == kYPlusB || MODE == kRes || MODE == kJac || MODE == kRowScaledRes) ob = b[row];
      if (MODE == kJac) ox = x[row];
      if (MODE == kYAdd) oy = y[row];
      if (MODE == kJac || MODE == kRowScaled || MODE == kRowScaledRes || MODE == kRestrictJ0) oa = (T)aux[row];
    }
    T sum = T(0);
    if (active) {
      const int e1 = rp[row + 1];
#pragma unroll 4
      for (int e = rp[row] + g; e < e1; e += G) sum += v[e] * x[ci[e]];
    }
    if (G > 1) sum = group_sum<G>(sum);
    if (active && g == 0) {
      if (MODE == kY) y[row] = sum;
      else if (MODE == kYAdd) y[row] = oy + sum;
      else if (MODE == kYPlusB) y[row] = ob + sum;
      else if (MODE == kRes) y[row] = ob - sum;
      else if (MODE == kJac) y[row] = ox + w * oa * (ob - sum);
      else if (MODE == kRowScaled) y[row] = oa * sum;
      else if (MODE == kRowScaledRes) y[row] = oa * (ob - sum);
      else if (MODE == kRestrictJ0) {
        y[row] = sum;
        y2[row] = w * oa * sum;
      }
    }
  }
}

// Padded FP32 CSR (Csr4F32): G lanes per row, each lane 4 entries per 16-byte load.
template <int G, int MODE, bool OFF16>
__global__ __launch_bounds__(kBlock) void k_csr4(int n, const int* __restrict__ rp4, const int4* __restrict__ ci4,
                                                  const short4* __restrict__ co4, const float4* __restrict__ v4,
                                                  const float* __restrict__ x, const float* __restrict__ b,
                                                  const float* __restrict__ aux, float w, float* __restrict__ y) {
  constexpr int RPB = kBlock / G;
  const int g = threadIdx.x % G;
  for (long long base = xcd_block() * RPB; base < n; base += (long long)gridDim.x * RPB) {
    const long long row = base + threadIdx.x / G;
    const bool active = row < n;
    // the row's own operands are loaded before the row walk (as in k_csr)
    float ob = 0.f, ox = 0.f, oa = 0.f;
    if (active && g == 0) {
      ob = b[row];
      if (MODE == kJac) {
        ox = x[row];
        oa = aux[row];
      }
    }
    float sum = 0.f;
    if (active) {
      const int q1 = rp4[row + 1] >> 2;
#pragma unroll 2
      for (int q = (rp4[row] >> 2) + g; q < q1; q += G) {
        const float4 a = v4[q];
        int4 c;
        if (OFF16) {
          const short4 o = co4[q];
          const int r = (int)row;
          c = make_int4(r + o.x, r + o.y, r + o.z, r + o.w);
        } else {
          c = ci4[q];
        }
        sum += a.x * x[c.x];
        sum += a.y * x[c.y];
        sum += a.z * x[c.z];
        sum += a.w * x[c.w];
      }
    }
    if (G > 1) sum = group_sum<G>(sum);
    if (active && g == 0) {
      if (MODE == kRes) y[row] = ob - sum;
      else y[row] = ox + w * oa * (ob - sum);
    }
  }
}

// Padded FP32 CSR over wave tiles (Csr4F32::wrow/wq): a wave takes one tile of
// whole rows, one 4-entry chunk per lane, so no lane idles on a short row. A lane
// finds its row by a binary search (shuffles) over the start chunks of the
// tile's rows, sums its chunk, and the lanes of a row add up by a segmented
// scan; the row's last lane applies the update. The sweeps are bound by each
// wave's chain of dependent loads at full occupancy, so the number of waves sets
// their time: on the 1M-cell hierarchy (4.6 chunks per row on level 1) G = 8
// lanes per row left 43% of the lanes idle; the tiles took a step from
// 8.25-8.27 to 7.86-7.87 ms.
template <int MODE, bool OFF16>
__global__ __launch_bounds__(kBlock) void k_csr4_tile(int nw, const int* __restrict__ wrow, const int* __restrict__ wq,
                                                       const int* __restrict__ rp4, const int4* __restrict__ ci4,
                                                       const short4* __restrict__ co4, const float4* __restrict__ v4,
                                                       const float* __restrict__ x, const float* __restrict__ b,
                                                       const float* __restrict__ aux, float w, float* __restrict__ y,
                                                       float* __restrict__ y2) {
  constexpr int WPB = kBlock / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  for (long long t = xcd_block() * WPB + threadIdx.x / kWave; t < nw; t += (long long)gridDim.x * WPB) {
    const int r0 = wrow[t], nr = wrow[t + 1] - r0;
    const int q0 = wq[t], q1 = wq[t + 1];
    const int q = q0 + lane;
    const bool valid = q < q1;
    // start chunk, relative to the tile, of row r0 + 1 + lane (64: no such row)
    const int ps = (lane + 1 < nr) ? (rp4[r0 + 1 + lane] >> 2) - q0 : kWave;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    short4 o = make_short4(0, 0, 0, 0);
    int4 c4 = make_int4(0, 0, 0, 0);
    if (valid) {
      a = v4[q];
      if (OFF16) o = co4[q];
      else c4 = ci4[q];
    }
    // rows of the tile that start at or before this lane's chunk
    int lo = 0;
#pragma unroll
    for (int st = kWave / 2; st >= 1; st >>= 1) {
      const int pc = __shfl(ps, lo + st - 1, kWave);
      if (pc <= lane) lo += st;
    }
    const int row = r0 + lo;
    const int rn = __shfl_down(row, 1, kWave);
    const bool last = valid && (lane == kWave - 1 || q + 1 >= q1 || rn != row);
    float ob = 0.f, ox = 0.f, oa = 0.f;
    if (last) {
      if (MODE == kRes || MODE == kJac) ob = b[row];
      if (MODE == kJac) ox = x[row];
      if (MODE == kJac || MODE == kRestrictJ0) oa = aux[row];
    }
    float sum = 0.f;
    if (valid) {
      int4 c;
      if (OFF16) c = make_int4(row + o.x, row + o.y, row + o.z, row + o.w);
      else c = c4;
      sum += a.x * x[c.x];
      sum += a.y * x[c.y];
      sum += a.z * x[c.z];
      sum += a.w * x[c.w];
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const float u = __shfl_up(sum, d, kWave);
      const int ru = __shfl_up(row, d, kWave);
      if (lane >= d && ru == row) sum += u;
    }
    if (last) {
      if (MODE == kRes) y[row] = ob - sum;
      else if (MODE == kJac) y[row] = ox + w * oa * (ob - sum);
      else if (MODE == kY) y[row] = sum;
      else if (MODE == kRestrictJ0) {
        y[row] = sum;
        y2[row] = w * oa * sum;
      }
    }
  }
}

template <int MODE, bool OFF16>
void csr4_tile_launch(const Csr4F32& A, const float* x, const float* b, const float* aux, float w, float* y,
                      hipStream_t s, float* y2 = nullptr) {
  const int grid = grid_for(A.nw, kBlock / kWave, 1 << 18);
  klaunch(k_csr4_tile<MODE, OFF16>, dim3(grid), dim3(kBlock), 0, s, A.nw, A.wrow, A.wq, A.rp4, A.ci4, A.co4, A.v4, x, b,
          aux, w, y, y2);
}

template <int MODE, bool OFF16>
void csr4_launch(const Csr4F32& A, int G, const float* x, const float* b, const float* aux, float w, float* y,
                 hipStream_t s) {
  const int grid = grid_for(A.n, kBlock / G, 1 << 18);
#define IMPES_CSR4_CASE(GG)                                                                                     \
  case GG:                                                                                                     \
    klaunch(k_csr4<GG, MODE, OFF16>, dim3(grid), dim3(kBlock), 0, s, A.n, A.rp4, A.ci4, A.co4, A.v4, x, b, aux, \
            w, y);                                                                                             \
    break;
  switch (G) {
    IMPES_CSR4_CASE(1)
    IMPES_CSR4_CASE(2)
    IMPES_CSR4_CASE(4)
    IMPES_CSR4_CASE(8)
    IMPES_CSR4_CASE(16)
    IMPES_CSR4_CASE(32)
    default:
      klaunch(k_csr4<64, MODE, OFF16>, dim3(grid), dim3(kBlock), 0, s, A.n, A.rp4, A.ci4, A.co4, A.v4, x, b, aux,
              w, y);
      break;
  }
#undef IMPES_CSR4_CASE
  IMPES_LAUNCH_CHECK();
}

template <int MODE>
void csr4_dispatch(const Csr4F32& A, int G, const float* x, const float* b, const float* aux, float w, float* y,
                   hipStream_t s) {
  if (A.n <= 0) return;
  IMPES_REQUIRE(((uintptr_t)A.ci4 & 15) == 0 && ((uintptr_t)A.v4 & 15) == 0, "padded CSR: arrays not 16-byte aligned");
  if (A.co4) IMPES_REQUIRE(((uintptr_t)A.co4 & 7) == 0, "padded CSR: offsets not 8-byte aligned");
  if (A.wrow) {
    if (A.co4) csr4_tile_launch<MODE, true>(A, x, b, aux, w, y, s);
    else csr4_tile_launch<MODE, false>(A, x, b, aux, w, y, s);
  } else if (A.co4) {
    csr4_launch<MODE, true>(A, G, x, b, aux, w, y, s);
  } else {
    csr4_launch<MODE, false>(A, G, x, b, aux, w, y, s);
  }
}

__global__ void k_csr4_count(int n, const int* __restrict__ rp, int* __restrict__ cnt) {
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < n; r += (long long)gridDim.x * blockDim.x)
    cnt[r] = max(4, (rp[r + 1] - rp[r] + 3) & ~3);
}

__global__ void k_csr4_pattern(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const int* __restrict__ rp4, int* __restrict__ ci4, int square) {
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < n; r += (long long)gridDim.x * blockDim.x) {
    const int e0 = rp[r], len = rp[r + 1] - e0, q0 = rp4[r], len4 = rp4[r + 1] - q0;
    const int pad = square ? (int)r : (len > 0 ? ci[e0] : 0);
    for (int k = 0; k < len4; ++k) ci4[q0 + k] = (k < len) ? ci[e0 + k] : pad;
  }
}

__global__ void k_csr4_offsets(int n, const int* __restrict__ rp4, const int* __restrict__ ci4,
                               short* __restrict__ co4, int* __restrict__ maxabs) {
  int m = 0;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < n; r += (long long)gridDim.x * blockDim.x) {
    for (int e = rp4[r]; e < rp4[r + 1]; ++e) {
      const int d = ci4[e] - (int)r;
      m = max(m, abs(d));
      co4[e] = (short)max(-32768, min(32767, d));
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & (kWave - 1)) == 0 && m > 0) atomicMax(maxabs, m);
}

// G lanes per row (entries strided by G): consecutive lanes write consecutive entries
template <int G>
__global__ void k_csr4_values(int n, const int* __restrict__ rp, const double* __restrict__ v,
                              const int* __restrict__ rp4, float* __restrict__ v4) {
  const int g = threadIdx.x % G;
  for (long long r = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / G; r < n;
       r += (long long)gridDim.x * blockDim.x / G) {
    const int e0 = rp[r], len = rp[r + 1] - e0, q0 = rp4[r], len4 = rp4[r + 1] - q0;
    for (int k = g; k < len4; k += G) v4[q0 + k] = (k < len) ? (float)v[e0 + k] : 0.f;
  }
}

template <class T, class TA, int MODE>
void csr_dispatch(int n, const int* rp, const int* ci, const T* v, const T* x, const T* b, const TA* aux, T w, T* y,
                  int G, hipStream_t s, T* y2 = nullptr) {
  if (n <= 0) return;
  const int rpb = kBlock / G;
  const int grid = grid_for(n, rpb, 1 << 18);
  switch (G) {
    case 1: klaunch(k_csr<1, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    case 2: klaunch(k_csr<2, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    case 4: klaunch(k_csr<4, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    case 8: klaunch(k_csr<8, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    case 16: klaunch(k_csr<16, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    case 32: klaunch(k_csr<32, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
    default: klaunch(k_csr<64, T, TA, MODE>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, b, aux, w, y, y2); break;
  }
  IMPES_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- stencil 7
template <class T>
struct St7Row {
  // sum_j A[c,j] x[j] in the column order of the CSR form (c-sx, c-nz, c-1, c, c+1, c+nz, c+sx)
  __device__ __forceinline__ static T apply(const Stencil7<T>& S, const T* __restrict__ x, long long c, int i, int j,
                                            int k) {
    const long long sy = S.nz;
    const long long sx = (long long)S.ny * S.nz;
    T s = T(0);
    if (i > 0) s += S.ox[c - sx] * x[c - sx];
    if (j > 0) s += S.oy[c - sy] * x[c - sy];
    if (k > 0) s += S.oz[c - 1] * x[c - 1];
    s += S.d[c] * x[c];
    if (k < S.nz - 1) s += S.oz[c] * x[c + 1];
    if (j < S.ny - 1) s += S.oy[c] * x[c + sy];
    if (i < S.nx - 1) s += S.ox[c] * x[c + sx];
    return s;
  }
  // the same sum with x[q] given by xv(q)
  template <class F>
  __device__ __forceinline__ static T apply_f(const Stencil7<T>& S, F xv, long long c, int i, int j, int k) {
    const long long sy = S.nz;
    const long long sx = (long long)S.ny * S.nz;
    T s = T(0);
    if (i > 0) s += S.ox[c - sx] * xv(c - sx);
    if (j > 0) s += S.oy[c - sy] * xv(c - sy);
    if (k > 0) s += S.oz[c - 1] * xv(c - 1);
    s += S.d[c] * xv(c);
    if (k < S.nz - 1) s += S.oz[c] * xv(c + 1);
    if (j < S.ny - 1) s += S.oy[c] * xv(c + sy);
    if (i < S.nx - 1) s += S.ox[c] * xv(c + sx);
    return s;
  }
};

template <class T, class TA, int MODE>
__global__ __launch_bounds__(kBlock) void k_st7(Stencil7<T> S, const T* __restrict__ x, const T* __restrict__ b,
                                                const TA* __restrict__ aux, T w, T* __restrict__ y) {
  const long long n = (long long)S.nx * S.ny * S.nz;
  for (long long c = xcd_block() * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(c % S.nz);
    const long long t = c / S.nz;
    const int j = (int)(t % S.ny);
    const int i = (int)(t / S.ny);
    if constexpr (MODE == kJacZ0) {
      auto xv = [&](long long q) { return w * (T)aux[q] * b[q]; };
      const T sum = St7Row<T>::apply_f(S, xv, c, i, j, k);
      y[c] = xv(c) + w * (T)aux[c] * (b[c] - sum);
      continue;
    }
    const T sum = St7Row<T>::apply(S, x, c, i, j, k);
    if (MODE == kY) y[c] = sum;
    else if (MODE == kRes) y[c] = b[c] - sum;
    else if (MODE == kJac) y[c] = x[c] + w * (T)aux[c] * (b[c] - sum);
    else if (MODE == kRowScaled) y[c] = (T)aux[c] * sum;
    else if (MODE == kRowScaledRes) y[c] = (T)aux[c] * (b[c] - sum);
  }
}

// Wide form of k_st7 for nz a multiple of V = 16 / sizeof(T): a thread takes V
// consecutive cells of one z-column (same i, j) and reads every coefficient and
// x array with 16-byte loads (the +-x and +-y neighbours of V aligned cells are
// V aligned cells), plus oz[c-1], x[c-1] and x[c+V] for the column ends. Each
// cell's sum is formed in St7Row's order, so the results equal k_st7's.
template <class T>
struct Wide;
template <>
struct Wide<float> { using type = float4; };
template <>
struct Wide<double> { using type = double2; };

template <class T, int V>
__device__ __forceinline__ void ld_wide(const T* __restrict__ p, T (&a)[V]) {
  static_assert(V * sizeof(T) == 16, "16-byte vectors");
  const typename Wide<T>::type v = *reinterpret_cast<const typename Wide<T>::type*>(p);
  __builtin_memcpy(a, &v, 16);
}
template <class T, int V>
__device__ __forceinline__ void st_wide(T* __restrict__ p, const T (&a)[V]) {
  typename Wide<T>::type v;
  __builtin_memcpy(&v, a, 16);
  *reinterpret_cast<typename Wide<T>::type*>(p) = v;
}

template <class T, int MODE>
__global__ __launch_bounds__(kBlock) void k_st7_wide(Stencil7<T> S, const T* __restrict__ x, const T* __restrict__ b,
                                                     const T* __restrict__ aux, T w, T* __restrict__ y) {
  constexpr int V = 16 / sizeof(T);
  const long long n = (long long)S.nx * S.ny * S.nz;
  const long long sy = S.nz;
  const long long sx = (long long)S.ny * S.nz;
  for (long long q = xcd_block() * blockDim.x + threadIdx.x; q * V < n; q += (long long)gridDim.x * blockDim.x) {
    const long long c = q * V;
    const int k0 = (int)(c % S.nz);
    const long long t = c / S.nz;
    const int j = (int)(t % S.ny);
    const int i = (int)(t / S.ny);
    T xc[V], dd[V], fz[V], fy[V], fx[V];
    T fw[V] = {}, xw[V] = {}, fs[V] = {}, xs[V] = {}, xn[V] = {}, xe[V] = {}, bb[V] = {}, aa[V] = {};
    // x at V aligned cells: read, or (kJacZ0) the first sweep w * aux .* b formed in place
    auto xg = [&](long long q, T (&dst)[V]) {
      if constexpr (MODE == kJacZ0) {
        T bq[V], aq[V];
        ld_wide(b + q, bq);
        ld_wide(aux + q, aq);
#pragma unroll
        for (int m = 0; m < V; ++m) dst[m] = w * aq[m] * bq[m];
      } else {
        ld_wide(x + q, dst);
      }
    };
    auto x1 = [&](long long q) {
      if constexpr (MODE == kJacZ0) return w * aux[q] * b[q];
      else return x[q];
    };
    xg(c, xc);
    ld_wide(S.d + c, dd);
    ld_wide(S.oz + c, fz);
    ld_wide(S.oy + c, fy);
    ld_wide(S.ox + c, fx);
    if (i > 0) { ld_wide(S.ox + c - sx, fw); xg(c - sx, xw); }
    if (j > 0) { ld_wide(S.oy + c - sy, fs); xg(c - sy, xs); }
    if (j < S.ny - 1) xg(c + sy, xn);
    if (i < S.nx - 1) xg(c + sx, xe);
    const T zb = k0 > 0 ? S.oz[c - 1] : T(0);
    const T xb = k0 > 0 ? x1(c - 1) : T(0);
    const T xa = k0 + V < S.nz ? x1(c + V) : T(0);
    if (MODE == kRes || MODE == kJac || MODE == kRowScaledRes || MODE == kJacZ0) ld_wide(b + c, bb);
    if (MODE == kJac || MODE == kRowScaled || MODE == kRowScaledRes || MODE == kJacZ0) ld_wide(aux + c, aa);
    T out[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int k = k0 + m;
      T sum = T(0);
      if (i > 0) sum += fw[m] * xw[m];
      if (j > 0) sum += fs[m] * xs[m];
      if (k > 0) sum += (m == 0 ? zb : fz[m > 0 ? m - 1 : 0]) * (m == 0 ? xb : xc[m > 0 ? m - 1 : 0]);
      sum += dd[m] * xc[m];
      if (k < S.nz - 1) sum += fz[m] * (m == V - 1 ? xa : xc[m < V - 1 ? m + 1 : 0]);
      if (j < S.ny - 1) sum += fy[m] * xn[m];
      if (i < S.nx - 1) sum += fx[m] * xe[m];
      if (MODE == kY) out[m] = sum;
      else if (MODE == kRes) out[m] = bb[m] - sum;
      else if (MODE == kJac || MODE == kJacZ0) out[m] = xc[m] + w * aa[m] * (bb[m] - sum);
      else if (MODE == kRowScaled) out[m] = aa[m] * sum;
      else out[m] = aa[m] * (bb[m] - sum);
    }
    st_wide(y + c, out);
  }
}

// y = dinv .* (A z) and z itself, with z = sout .* x formed on the fly from the
// V-cycle's output x (FP32 or FP64): the preconditioner's output scaling fused
// into the Krylov operator apply. Same operation order as k_scale followed by
// k_st7<kRowScaled>, so the result is identical to the two-kernel form.
template <class TX>
__global__ __launch_bounds__(kBlock) void k_st7_rowscaled_z(Stencil7<double> S, const double* __restrict__ dinv,
                                                            const double* __restrict__ sout,
                                                            const TX* __restrict__ x, double* __restrict__ z,
                                                            double* __restrict__ y) {
  const long long n = (long long)S.nx * S.ny * S.nz;
  const long long sy = S.nz;
  const long long sx = (long long)S.ny * S.nz;
  for (long long c = xcd_block() * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(c % S.nz);
    const long long t = c / S.nz;
    const int j = (int)(t % S.ny);
    const int i = (int)(t / S.ny);
    auto zv = [&](long long q) { return sout[q] * (double)x[q]; };
    const double zc = zv(c);
    double s = 0.0;
    if (i > 0) s += S.ox[c - sx] * zv(c - sx);
    if (j > 0) s += S.oy[c - sy] * zv(c - sy);
    if (k > 0) s += S.oz[c - 1] * zv(c - 1);
    s += S.d[c] * zc;
    if (k < S.nz - 1) s += S.oz[c] * zv(c + 1);
    if (j < S.ny - 1) s += S.oy[c] * zv(c + sy);
    if (i < S.nx - 1) s += S.ox[c] * zv(c + sx);
    z[c] = zc;
    y[c] = dinv[c] * s;
  }
}

// Wide form of k_st7_rowscaled_z for even nz: two cells of one z-column per
// thread, 16-byte loads of every FP64 array (8-byte loads of an FP32 x), the same
// operation order per cell.
template <class TX>
struct Pair;
template <>
struct Pair<float> { using type = float2; };
template <>
struct Pair<double> { using type = double2; };

template <class TX>
__device__ __forceinline__ void ld_pair(const TX* __restrict__ p, TX (&a)[2]) {
  const typename Pair<TX>::type v = *reinterpret_cast<const typename Pair<TX>::type*>(p);
  a[0] = v.x;
  a[1] = v.y;
}

template <class TX>
__global__ __launch_bounds__(kBlock) void k_st7_rowscaled_z_wide(Stencil7<double> S, const double* __restrict__ dinv,
                                                                 const double* __restrict__ sout,
                                                                 const TX* __restrict__ x, double* __restrict__ z,
                                                                 double* __restrict__ y) {
  const long long n = (long long)S.nx * S.ny * S.nz;
  const long long sy = S.nz;
  const long long sx = (long long)S.ny * S.nz;
  for (long long q = xcd_block() * blockDim.x + threadIdx.x; q * 2 < n; q += (long long)gridDim.x * blockDim.x) {
    const long long c = q * 2;
    const int k0 = (int)(c % S.nz);
    const long long t = c / S.nz;
    const int j = (int)(t % S.ny);
    const int i = (int)(t / S.ny);
    // z = sout .* x at the cell pair and at each neighbour pair
    auto zpair = [&](long long p0, double (&zz)[2]) {
      TX xx[2];
      double ss[2];
      ld_pair(x + p0, xx);
      ld_pair(sout + p0, ss);
      zz[0] = ss[0] * (double)xx[0];
      zz[1] = ss[1] * (double)xx[1];
    };
    double zc[2], dd[2], fz[2], fy[2], fx[2], di[2];
    double zw[2] = {}, fw[2] = {}, zs[2] = {}, fs[2] = {}, zn[2] = {}, ze[2] = {};
    zpair(c, zc);
    ld_pair(S.d + c, dd);
    ld_pair(S.oz + c, fz);
    ld_pair(S.oy + c, fy);
    ld_pair(S.ox + c, fx);
    ld_pair(dinv + c, di);
    if (i > 0) { ld_pair(S.ox + c - sx, fw); zpair(c - sx, zw); }
    if (j > 0) { ld_pair(S.oy + c - sy, fs); zpair(c - sy, zs); }
    if (j < S.ny - 1) zpair(c + sy, zn);
    if (i < S.nx - 1) zpair(c + sx, ze);
    const double fb = k0 > 0 ? S.oz[c - 1] : 0.0;
    const double zb = k0 > 0 ? sout[c - 1] * (double)x[c - 1] : 0.0;
    const double za = k0 + 2 < S.nz ? sout[c + 2] * (double)x[c + 2] : 0.0;
    double out[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int k = k0 + m;
      double sum = 0.0;
      if (i > 0) sum += fw[m] * zw[m];
      if (j > 0) sum += fs[m] * zs[m];
      if (k > 0) sum += (m == 0 ? fb : fz[0]) * (m == 0 ? zb : zc[0]);
      sum += dd[m] * zc[m];
      if (k < S.nz - 1) sum += fz[m] * (m == 1 ? za : zc[1]);
      if (j < S.ny - 1) sum += fy[m] * zn[m];
      if (i < S.nx - 1) sum += fx[m] * ze[m];
      out[m] = di[m] * sum;
    }
    *reinterpret_cast<double2*>(z + c) = make_double2(zc[0], zc[1]);
    *reinterpret_cast<double2*>(y + c) = make_double2(out[0], out[1]);
  }
}

template <class T>
__global__ void k_jacobi0(const T* __restrict__ b, const T* __restrict__ inv_relax, T w, T* __restrict__ xo, int n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    xo[i] = w * inv_relax[i] * b[i];
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <class T, class TA, int MODE>
void st7_launch(const Stencil7<T>& S, const T* x, const T* b, const TA* aux, T w, T* y, hipStream_t s) {
  const long long n = (long long)S.nx * S.ny * S.nz;
  if (n <= 0) return;
  constexpr int V = 16 / sizeof(T);
  constexpr bool uses_b = MODE == kRes || MODE == kJac || MODE == kRowScaledRes || MODE == kJacZ0;
  constexpr bool uses_aux = MODE == kJac || MODE == kRowScaled || MODE == kRowScaledRes || MODE == kJacZ0;
  if constexpr (std::is_same<T, TA>::value) {
    // the wide form needs whole 16-byte vectors along z and 16-byte aligned arrays
    if (S.nz % V == 0 && aligned16(x) && aligned16(y) && aligned16(S.d) && aligned16(S.ox) && aligned16(S.oy) &&
        aligned16(S.oz) && (!uses_b || aligned16(b)) && (!uses_aux || aligned16(aux))) {
      klaunch(k_st7_wide<T, MODE>, dim3(grid_for(n / V, kBlock, 1 << 18)), dim3(kBlock), 0, s, S, x, b, aux, w, y);
      IMPES_LAUNCH_CHECK();
      return;
    }
  }
  klaunch(k_st7<T, TA, MODE>, dim3(grid_for(n, kBlock, 1 << 18)), dim3(kBlock), 0, s, S, x, b, aux, w, y);
  IMPES_LAUNCH_CHECK();
}

}  // namespace

// Algorithmic HBM bytes of one pass (each operand read or written once):
// CSR: row pointers + (column, value) per entry + `vecs` row vectors besides x
// (x read once, e.g. b, inv_relax and the output for Jacobi);
// stencil: four coefficient arrays + x + `vecs` row vectors.
double csr_bytes(long long n, long long nnz, int val_bytes, int vecs) {
  return (double)(n + 1) * 4 + (double)nnz * (4 + val_bytes) + (double)n * val_bytes * (1 + vecs);
}
template <class T>
double st7_bytes(const Stencil7<T>& S, int val_bytes, int vecs) {
  const double n = (double)S.nx * S.ny * S.nz;
  return n * val_bytes * (4 + 1 + vecs);
}

// Lanes per row: the smallest power of two with at most ~4 entries per lane, so a
// lane has several independent (value, column, gather) loads in flight (one entry
// per lane is latency-bound). A sweep too small to fill the GPU (fewer than 2^18
// lanes) is bound by each row's serial chain instead of by throughput: it takes
// twice the lanes (~2 entries per lane). Measured at 1M cells: level-1/2 sweeps
// 21.9/23.6 us at ~4 per lane against 25.5/28.4 us at ~2; level-4+ sweeps and
// coarse restrictions 7-12 us at ~4 against 5-8 us at ~2.
int csr_group_for(int n_rows, int nnz) {
  if (n_rows <= 0) return 1;
  const double mean = (double)nnz / (double)n_rows;
  int g = 1;
  while (g < 64 && g * 4 < mean) g <<= 1;
  if ((long long)n_rows * g < (1LL << 18) && g < 64 && g * 2 < mean) g <<= 1;
  return g;
}

void launch_spmv_f64(const CsrF64& A, const double* x, double* y, int G, hipStream_t s) {
  csr_dispatch<double, double, kY>(A.n_rows, A.rp, A.ci, A.v, x, nullptr, (const double*)nullptr, 0.0, y, G, s);
}
void launch_spmv_f32(const CsrF32& A, const float* x, float* y, int G, hipStream_t s) {
  csr_dispatch<float, float, kY>(A.n_rows, A.rp, A.ci, A.v, x, nullptr, (const float*)nullptr, 0.f, y, G, s);
}
void launch_spmv_add_f64(const CsrF64& A, const double* x, double* y, int G, hipStream_t s) {
  csr_dispatch<double, double, kYAdd>(A.n_rows, A.rp, A.ci, A.v, x, nullptr, (const double*)nullptr, 0.0, y, G, s);
}
void launch_spmv_add_f32(const CsrF32& A, const float* x, float* y, int G, hipStream_t s) {
  csr_dispatch<float, float, kYAdd>(A.n_rows, A.rp, A.ci, A.v, x, nullptr, (const float*)nullptr, 0.f, y, G, s);
}
void launch_spmv_add_to_f64(const CsrF64& A, const double* x, const double* y0, double* y, int G, hipStream_t s) {
  csr_dispatch<double, double, kYPlusB>(A.n_rows, A.rp, A.ci, A.v, x, y0, (const double*)nullptr, 0.0, y, G, s);
}
void launch_spmv_add_to_f32(const CsrF32& A, const float* x, const float* y0, float* y, int G, hipStream_t s) {
  csr_dispatch<float, float, kYPlusB>(A.n_rows, A.rp, A.ci, A.v, x, y0, (const float*)nullptr, 0.f, y, G, s);
}
void launch_residual_f64(const CsrF64& A, const double* x, const double* b, double* r, int G, hipStream_t s) {
  csr_dispatch<double, double, kRes>(A.n_rows, A.rp, A.ci, A.v, x, b, (const double*)nullptr, 0.0, r, G, s);
}
void launch_residual_f32(const CsrF32& A, const float* x, const float* b, float* r, int G, hipStream_t s) {
  KScope ks("csr_residual_f32", csr_bytes(A.n_rows, A.nnz, 4, 2), s);
  csr_dispatch<float, float, kRes>(A.n_rows, A.rp, A.ci, A.v, x, b, (const float*)nullptr, 0.f, r, G, s);
}
void launch_jacobi_f64(const CsrF64& A, const double* xi, const double* b, const double* inv_relax, double w,
                       double* xo, int G, hipStream_t s) {
  KScope ks("csr_jacobi_f64", csr_bytes(A.n_rows, A.nnz, 8, 3), s);
  csr_dispatch<double, double, kJac>(A.n_rows, A.rp, A.ci, A.v, xi, b, inv_relax, w, xo, G, s);
}
void launch_jacobi_f32(const CsrF32& A, const float* xi, const float* b, const float* inv_relax, float w, float* xo,
                       int G, hipStream_t s) {
  KScope ks("csr_jacobi_f32", csr_bytes(A.n_rows, A.nnz, 4, 3), s);
  csr_dispatch<float, float, kJac>(A.n_rows, A.rp, A.ci, A.v, xi, b, inv_relax, w, xo, G, s);
}
int csr4_wave_tiles(int n, const int* rp4, std::vector<int>& wrow, std::vector<int>& wq) {
  wrow.assign(1, 0);
  wq.assign(1, 0);
  int cur = 0;
  for (int r = 0; r < n; ++r) {
    const int len = (rp4[r + 1] - rp4[r]) >> 2;
    if (len > kWave) return 0;
    if (cur + len > kWave) {
      wrow.push_back(r);
      wq.push_back(rp4[r] >> 2);
      cur = 0;
    }
    cur += len;
  }
  wrow.push_back(n);
  wq.push_back(rp4[n] >> 2);
  return (int)wrow.size() - 1;
}
int csr4_group_for(int n_rows, long long nnz4) {
  if (n_rows <= 0) return 1;
  // at most one 4-entry chunk per lane on a row of mean length: the smallest power
  // of two not below the mean number of chunks (measured on the 1M-cell IMPES
  // hierarchy: 8 lanes on level 1 with ~4 chunks per row beat 4 lanes by 7%)
  const double chunks = (double)nnz4 / 4.0 / (double)n_rows;
  int g = 1;
  while (g < 64 && g < chunks) g <<= 1;
  return g;
}
void launch_csr4_count(int n, const int* rp, int* cnt, hipStream_t s) {
  if (n <= 0) return;
  k_csr4_count<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp, cnt);
  IMPES_LAUNCH_CHECK();
}
void launch_csr4_pattern(int n, const int* rp, const int* ci, const int* rp4, int* ci4, hipStream_t s, bool square) {
  if (n <= 0) return;
  k_csr4_pattern<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp, ci, rp4, ci4, square ? 1 : 0);
  IMPES_LAUNCH_CHECK();
}
void launch_csr4_values(int n, const int* rp, const double* v, const int* rp4, float* v4, hipStream_t s) {
  if (n <= 0) return;
  k_csr4_values<8><<<grid_for((long long)n * 8, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp, v, rp4, v4);
  IMPES_LAUNCH_CHECK();
}
void launch_csr4_offsets(int n, const int* rp4, const int* ci4, short* co4, int* maxabs, hipStream_t s) {
  if (n <= 0) return;
  k_csr4_offsets<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp4, ci4, co4, maxabs);
  IMPES_LAUNCH_CHECK();
}
void launch_residual4_f32(const Csr4F32& A, int G, const float* x, const float* b, float* r, hipStream_t s) {
  KScope ks("csr_residual_f32", csr_bytes(A.n, A.nnz, 4, 2), s);
  csr4_dispatch<kRes>(A, G, x, b, nullptr, 0.f, r, s);
}
void launch_jacobi4_f32(const Csr4F32& A, int G, const float* xi, const float* b, const float* inv_relax, float w,
                        float* xo, hipStream_t s) {
  KScope ks("csr_jacobi_f32", csr_bytes(A.n, A.nnz, 4, 3), s);
  csr4_dispatch<kJac>(A, G, xi, b, inv_relax, w, xo, s);
}
void launch_spmv4_f32(const Csr4F32& A, const float* x, float* y, hipStream_t s) {
  IMPES_REQUIRE(A.wrow && !A.co4, "padded transfer: wave tiles and column indices required");
  if (A.n <= 0) return;
  csr4_tile_launch<kY, false>(A, x, nullptr, nullptr, 0.f, y, s);
  IMPES_LAUNCH_CHECK();
}
void launch_restrict4_j0_f32(const Csr4F32& R, const float* r, const float* inv_relax_c, float w, float* bc, float* xc,
                             hipStream_t s) {
  IMPES_REQUIRE(R.wrow && !R.co4, "padded transfer: wave tiles and column indices required");
  if (R.n <= 0) return;
  csr4_tile_launch<kRestrictJ0, false>(R, r, nullptr, inv_relax_c, w, bc, s, xc);
  IMPES_LAUNCH_CHECK();
}
void launch_restrict_j0_f64(const CsrF64& R, const double* r, const double* inv_relax_c, double w, double* bc,
                            double* xc, int G, hipStream_t s) {
  csr_dispatch<double, double, kRestrictJ0>(R.n_rows, R.rp, R.ci, R.v, r, nullptr, inv_relax_c, w, bc, G, s, xc);
}
void launch_restrict_j0_f32(const CsrF32& R, const float* r, const float* inv_relax_c, float w, float* bc, float* xc,
                            int G, hipStream_t s) {
  csr_dispatch<float, float, kRestrictJ0>(R.n_rows, R.rp, R.ci, R.v, r, nullptr, inv_relax_c, w, bc, G, s, xc);
}
void launch_jacobi0_f64(const double* b, const double* inv_relax, double w, double* xo, int n, hipStream_t s) {
  if (n <= 0) return;
  k_jacobi0<double><<<grid_for(n, kBlock, 1 << 18), kBlock, 0, s>>>(b, inv_relax, w, xo, n);
  IMPES_LAUNCH_CHECK();
}
void launch_jacobi0_f32(const float* b, const float* inv_relax, float w, float* xo, int n, hipStream_t s) {
  if (n <= 0) return;
  k_jacobi0<float><<<grid_for(n, kBlock, 1 << 18), kBlock, 0, s>>>(b, inv_relax, w, xo, n);
  IMPES_LAUNCH_CHECK();
}
void launch_csr_rowscaled_spmv_f64(const CsrF64& A, const double* dinv, const double* x, double* y, int G,
                                   hipStream_t s) {
  csr_dispatch<double, double, kRowScaled>(A.n_rows, A.rp, A.ci, A.v, x, nullptr, dinv, 0.0, y, G, s);
}
void launch_csr_rowscaled_residual_f64(const CsrF64& A, const double* dinv, const double* x, const double* b,
                                       double* r, int G, hipStream_t s) {
  csr_dispatch<double, double, kRowScaledRes>(A.n_rows, A.rp, A.ci, A.v, x, b, dinv, 0.0, r, G, s);
}

void launch_st7_spmv_f64(const Stencil7<double>& S, const double* x, double* y, hipStream_t s) {
  st7_launch<double, double, kY>(S, x, nullptr, (const double*)nullptr, 0.0, y, s);
}
void launch_st7_residual_f64(const Stencil7<double>& S, const double* x, const double* b, double* r, hipStream_t s) {
  st7_launch<double, double, kRes>(S, x, b, (const double*)nullptr, 0.0, r, s);
}
void launch_st7_jacobi_z0_f64(const Stencil7<double>& S, const double* b, const double* inv_relax, double w,
                              double* xo, hipStream_t s) {
  KScope ks("st7_jacobi_z0_f64", st7_bytes(S, 8, 2), s);
  st7_launch<double, double, kJacZ0>(S, nullptr, b, inv_relax, w, xo, s);
}
void launch_st7_jacobi_z0_f32(const Stencil7<float>& S, const float* b, const float* inv_relax, float w, float* xo,
                              hipStream_t s) {
  KScope ks("st7_jacobi_z0_f32", st7_bytes(S, 4, 2), s);
  st7_launch<float, float, kJacZ0>(S, nullptr, b, inv_relax, w, xo, s);
}
void launch_st7_residual_f32(const Stencil7<float>& S, const float* x, const float* b, float* r, hipStream_t s) {
  KScope ks("st7_residual_f32", st7_bytes(S, 4, 2), s);
  st7_launch<float, float, kRes>(S, x, b, (const float*)nullptr, 0.f, r, s);
}
void launch_st7_jacobi_f64(const Stencil7<double>& S, const double* xi, const double* b, const double* inv_relax,
                           double w, double* xo, hipStream_t s) {
  KScope ks("st7_jacobi_f64", st7_bytes(S, 8, 3), s);
  st7_launch<double, double, kJac>(S, xi, b, inv_relax, w, xo, s);
}
void launch_st7_jacobi_f32(const Stencil7<float>& S, const float* xi, const float* b, const float* inv_relax, float w,
                           float* xo, hipStream_t s) {
  KScope ks("st7_jacobi_f32", st7_bytes(S, 4, 3), s);
  st7_launch<float, float, kJac>(S, xi, b, inv_relax, w, xo, s);
}
void launch_st7_rowscaled_spmv_f64(const Stencil7<double>& S, const double* dinv, const double* x, double* y,
                                   hipStream_t s) {
  KScope ks("st7_rowscaled_spmv_f64", st7_bytes(S, 8, 2), s);
  st7_launch<double, double, kRowScaled>(S, x, nullptr, dinv, 0.0, y, s);
}
template <class TX>
static void st7_rowscaled_z(const Stencil7<double>& S, const double* dinv, const double* sout, const TX* x, double* z,
                            double* y, hipStream_t s) {
  const long long n = (long long)S.nx * S.ny * S.nz;
  if (n <= 0) return;
  // x and sout, 4 coefficient arrays and dinv read; z and y written
  KScope ks(sizeof(TX) == 4 ? "st7_rowscaled_z_f32in" : "st7_rowscaled_z_f64in",
            (double)n * (sizeof(TX) + 8 * 8), s);
  if (S.nz % 2 == 0 && aligned16(dinv) && aligned16(sout) && aligned16(z) && aligned16(y) && aligned16(S.d) &&
      aligned16(S.ox) && aligned16(S.oy) && aligned16(S.oz) && ((uintptr_t)x % (2 * sizeof(TX))) == 0) {
    klaunch(k_st7_rowscaled_z_wide<TX>, dim3(grid_for(n / 2, kBlock, 1 << 18)), dim3(kBlock), 0, s, S, dinv, sout, x,
            z, y);
  } else {
    klaunch(k_st7_rowscaled_z<TX>, dim3(grid_for(n, kBlock, 1 << 18)), dim3(kBlock), 0, s, S, dinv, sout, x, z, y);
  }
  IMPES_LAUNCH_CHECK();
}
void launch_st7_rowscaled_spmv_z_f64(const Stencil7<double>& S, const double* dinv, const double* sout,
                                     const float* x, double* z, double* y, hipStream_t s) {
  st7_rowscaled_z<float>(S, dinv, sout, x, z, y, s);
}
void launch_st7_rowscaled_spmv_z_f64(const Stencil7<double>& S, const double* dinv, const double* sout,
                                     const double* x, double* z, double* y, hipStream_t s) {
  st7_rowscaled_z<double>(S, dinv, sout, x, z, y, s);
}
void launch_st7_rowscaled_residual_f64(const Stencil7<double>& S, const double* dinv, const double* x,
                                       const double* b, double* r, hipStream_t s) {
  st7_launch<double, double, kRowScaledRes>(S, x, b, dinv, 0.0, r, s);
}

}  // namespace impes
