// FP64 restarted FGMRES(m) with CGS2 orthogonalisation (lagged by one step where a
// preconditioner is used and m <= 32). The Hessenberg matrix, Givens rotations and
// the residual estimate live on the device; the host reads one scalar per Arnoldi
// step (written by the device into pinned host memory) to decide termination.
//
// One Arnoldi step with the AMG preconditioner is, besides the V-cycle:
//   st7/CSR operator apply (fused with the V-cycle's output scaling on stencils),
//   three Gram-Schmidt passes (dots; update + dots; update + norm + rotation),
//   and the normalisation of the new basis vector, fused with the V-cycle's
//   input scaling for the next step.
#include <cstdint>

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
  // lagged orthogonalisation: provisional column (m+1) and the update's 2 scalars;
  // its coefficients a and c reuse h() and h2()
  __host__ __device__ int hp() const { return sc() + 8; }
  __host__ __device__ int lag() const { return hp() + m + 1; }
  __host__ __device__ int size() const { return lag() + 2; }
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

// Workgroups of a Gram-Schmidt pass: each folds its rows' k products per column
// with wavefront shuffles, so more rows per thread amortise the folds (1024: 1%
// faster per time step than 2048 at 1M cells; 512 and 256 no better).
constexpr int kGsBlocks = 1024;
static_assert(kGsBlocks <= kMultiDotBlocks, "the partials workspace is sized for kMultiDotBlocks");

template <int PASS>
void launch_gs_pass(int k, double* w, const double* V, long long n, const double* h, double* part, hipStream_t s) {
  const int nb = kGsBlocks;
  if (k <= 4)
    k_gs_pass<4, PASS><<<nb, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else if (k <= 8)
    k_gs_pass<8, PASS><<<nb, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else if (k <= 16)
    k_gs_pass<16, PASS><<<nb, kGsThreads, 0, s>>>(k, w, V, n, h, part);
  else
    k_gs_pass<kMaxFusedGs, PASS><<<nb, kGsThreads, 0, s>>>(k, w, V, n, h, part);
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

// ---------------------------------------------------------------- lagged CGS2
// Arnoldi step j with the second (re)orthogonalisation of each basis vector
// delayed by one step, so that the step makes two passes over the basis where
// CGS2 makes three (and needs no separate normalisation). Entering step j:
// v_0..v_{j-1} final and orthonormal, u = V[:, j] the previous step's vector
// orthogonalised once (unnormalised), w = D^-1 A z_j with z_j = M u (the
// V-cycle ran on u itself: FGMRES needs A z_j expressed in the orthonormal basis,
// not z_j tied to a basis vector, and M is linear).
//   dots    a = V^T u, c = V^T w, s = u.u, d = u.w                   (one pass)
//   colA    nu = sqrt(s - a.a) (1 for j = 0: u = v_0 is final); column j-1 of
//           H is final: hp_{j-1} + a, nu (hp: its projections at step j-1);
//           its Givens rotation; provisional column j: c, e = (d - a.c)/nu
//   update  v_j = (u - V a)/nu, w1 = w - V c - v_j e -> V[:, j+1], the next
//           V-cycle input S .* w1, and ||w1||^2                       (one pass)
//   colB    the residual estimate of step j from the provisional column
//           (c, e, ||w1||) under the final rotations 0..j-1
// v_j is exactly CGS2's: u reorthogonalised against the final basis; column j
// of H is completed one step later (or by a closing dots + colA when the cycle
// stops after step j).
// P lanes share a row (P = 1 or 2), lane h of the pair taking columns t = h + P m:
// the accumulators (and coefficients) per thread fall from 2K to 2K/P doubles.
template <int P, class T>
__device__ __forceinline__ T pair_sum(T v) {  // sum over the lanes of one column set (xor offsets >= P)
#pragma unroll
  for (int off = kWave / 2; off >= P; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <int K, int P, int R>
__global__ __launch_bounds__(kGsThreads) void k_lag_dots(int k, const double* __restrict__ V, long long n,
                                                         const double* __restrict__ u, const double* __restrict__ w,
                                                         double* __restrict__ part) {
  constexpr int KP = K / P;
  constexpr int NC = 2 * K + 2;
  constexpr int NW = kGsThreads / kWave;
  __shared__ double lds[NW * NC];
  const int h = threadIdx.x % P;
  double aa[KP], cc[KP];
#pragma unroll
  for (int m = 0; m < KP; ++m) aa[m] = cc[m] = 0.0;
  double ss = 0.0, dd = 0.0;
  const long long rows_per_pass = (long long)gridDim.x * blockDim.x / P;
  if constexpr (R == 1) {
    for (long long i = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / P; i < n; i += rows_per_pass) {
      const double ui = u[i], wi = w[i];
#pragma unroll
      for (int m = 0; m < KP; ++m) {
        const int t = h + P * m;
        const double v = (t < k) ? V[(long long)t * n + i] : 0.0;
        aa[m] += v * ui;
        cc[m] += v * wi;
      }
      if (h == 0) {
        ss += ui * ui;
        dd += ui * wi;
      }
    }
  } else {  // rows 2 i2 and 2 i2 + 1 with 16-byte loads (n even, arrays 16-byte aligned)
    for (long long i2 = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / P; 2 * i2 < n; i2 += rows_per_pass) {
      const long long i = 2 * i2;
      const double2 u2 = *reinterpret_cast<const double2*>(u + i), w2 = *reinterpret_cast<const double2*>(w + i);
#pragma unroll
      for (int m = 0; m < KP; ++m) {
        const int t = h + P * m;
        double2 v = make_double2(0.0, 0.0);
        if (t < k) v = *reinterpret_cast<const double2*>(V + (long long)t * n + i);
        aa[m] += v.x * u2.x;
        aa[m] += v.y * u2.y;
        cc[m] += v.x * w2.x;
        cc[m] += v.y * w2.y;
      }
      if (h == 0) {
        ss += u2.x * u2.x;
        ss += u2.y * u2.y;
        dd += u2.x * w2.x;
        dd += u2.y * w2.y;
      }
    }
  }
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int m = 0; m < KP; ++m) {
    if (P * m < k) {
      const double x = pair_sum<P>(aa[m]), y = pair_sum<P>(cc[m]);
      const int t = lane + P * m;  // lanes 0..P-1 hold the sums of columns P m .. P m + P - 1
      if (lane < P && t < k) {
        lds[wid * NC + t] = x;
        lds[wid * NC + K + t] = y;
      }
    }
  }
  {
    const double x = wave_sum(ss), y = wave_sum(dd);
    if (lane == 0) {
      lds[wid * NC + 2 * K] = x;
      lds[wid * NC + 2 * K + 1] = y;
    }
  }
  __syncthreads();
  // partial columns: a_t -> t, c_t -> k + t, s -> 2k, d -> 2k + 1
  for (int q = threadIdx.x; q < 2 * k + 2; q += blockDim.x) {
    const int src = (q < k) ? q : (q < 2 * k ? K + (q - k) : 2 * K + (q - 2 * k));
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < NW; ++r) v += lds[r * NC + src];
    part[(long long)q * gridDim.x + blockIdx.x] = v;
  }
}

template <int K, int P, class TO, int R>
__global__ __launch_bounds__(kGsThreads) void k_lag_update(int k, double* __restrict__ V, long long n,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ ca,
                                                           const double* __restrict__ cc,
                                                           const double* __restrict__ lag,
                                                           const double* __restrict__ S, TO* __restrict__ bout,
                                                           double* __restrict__ part) {
  constexpr int KP = K / P;
  __shared__ double lds[kGsThreads / kWave];
  const int h = threadIdx.x % P;
  double a[KP], c[KP];
#pragma unroll
  for (int m = 0; m < KP; ++m) {
    const int t = h + P * m;
    a[m] = (t < k) ? ca[t] : 0.0;
    c[m] = (t < k) ? cc[t] : 0.0;
  }
  const double inv_nu = lag[0], e = lag[1];
  double* u = V + (long long)k * n;
  double* un = u + n;
  double acc = 0.0;
  const long long rows_per_pass = (long long)gridDim.x * blockDim.x / P;
  // every lane of a pair walks the same rows (the pair folds with one shuffle)
  if constexpr (R == 1) {
    for (long long i = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / P; i < n; i += rows_per_pass) {
      double su = 0.0, sw = 0.0;
#pragma unroll
      for (int m = 0; m < KP; ++m) {
        const int t = h + P * m;
        if (t < k) {
          const double v = V[(long long)t * n + i];
          su += v * a[m];
          sw += v * c[m];
        }
      }
      if (P > 1) {
        su += __shfl_xor(su, 1, kWave);
        sw += __shfl_xor(sw, 1, kWave);
      }
      if (h == 0) {
        const double vj = (u[i] - su) * inv_nu;
        const double w1 = (w[i] - sw) - vj * e;
        u[i] = vj;
        un[i] = w1;
        bout[i] = (TO)(S[i] * w1);
        acc += w1 * w1;
      }
    }
  } else {  // rows 2 i2 and 2 i2 + 1 with 16-byte loads; each row's arithmetic as above
    for (long long i2 = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / P; 2 * i2 < n; i2 += rows_per_pass) {
      const long long i = 2 * i2;
      double sux = 0.0, suy = 0.0, swx = 0.0, swy = 0.0;
#pragma unroll
      for (int m = 0; m < KP; ++m) {
        const int t = h + P * m;
        if (t < k) {
          const double2 v = *reinterpret_cast<const double2*>(V + (long long)t * n + i);
          sux += v.x * a[m];
          suy += v.y * a[m];
          swx += v.x * c[m];
          swy += v.y * c[m];
        }
      }
      if (P > 1) {
        sux += __shfl_xor(sux, 1, kWave);
        suy += __shfl_xor(suy, 1, kWave);
        swx += __shfl_xor(swx, 1, kWave);
        swy += __shfl_xor(swy, 1, kWave);
      }
      if (h == 0) {
        const double2 u2 = *reinterpret_cast<const double2*>(u + i);
        const double2 w2 = *reinterpret_cast<const double2*>(w + i);
        const double2 s2 = *reinterpret_cast<const double2*>(S + i);
        const double vjx = (u2.x - sux) * inv_nu, vjy = (u2.y - suy) * inv_nu;
        const double w1x = (w2.x - swx) - vjx * e, w1y = (w2.y - swy) - vjy * e;
        *reinterpret_cast<double2*>(u + i) = make_double2(vjx, vjy);
        *reinterpret_cast<double2*>(un + i) = make_double2(w1x, w1y);
        bout[i] = (TO)(s2.x * w1x);
        bout[i + 1] = (TO)(s2.y * w1y);
        acc += w1x * w1x;
        acc += w1y * w1y;
      }
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// c, s = the Givens rotation that zeroes b under a
__device__ __forceinline__ void givens(double a, double b, double& c, double& s) {
  if (b == 0.0) { c = 1.0; s = 0.0; }
  else if (fabs(b) > fabs(a)) { const double t = a / b; s = 1.0 / sqrt(1.0 + t * t); c = s * t; }
  else { const double t = b / a; c = 1.0 / sqrt(1.0 + t * t); s = c * t; }
}

// All 2j + 2 columns of partials fold at once: G lanes per column (the widest
// power of two up to a wave with G * columns <= 1024 threads), each lane summing
// every G-th partial, then the G lanes. The partials were written by other XCDs,
// so each load is a far round trip: a lane keeps 16 of them in flight.
constexpr int kColAThreads = 1024;
static_assert(kColAThreads / 8 >= 2 * kMaxFusedGs + 2, "one lane group per column");
template <int G>
__device__ __forceinline__ void fold_cols(const double* __restrict__ part, int nb, int ncols, double* val) {
  const int q = threadIdx.x / G, g = threadIdx.x % G;
  double v = 0.0;
  if (q < ncols) {
#pragma unroll 16
    for (int b = g; b < nb; b += G) v += part[(long long)q * nb + b];
  }
  v = group_sum<G>(v);
  if (q < ncols && g == 0) val[q] = v;
}

__global__ __launch_bounds__(kColAThreads) void k_lag_colA(Layout L, int j, double* __restrict__ sm,
                                                           const double* __restrict__ part, int nb) {
  __shared__ double val[2 * kMaxFusedGs + 2];
  __shared__ double rot[3 * kMaxFusedGs];  // hp, cs, sn of rows 0..j-1 as they were on entry
  const int t = threadIdx.x;
  double* hp = sm + L.hp();
  double* g = sm + L.g();
  double* cs = sm + L.cs();
  double* sn = sm + L.sn();
  // the serial rotation below must not wait on far loads: fetch its rows now,
  // while the fold's loads are in flight
  double rh = 0.0, rc = 0.0, rs = 0.0;
  if (t < j) {
    rh = hp[t];
    rc = cs[t];
    rs = sn[t];
  }
  const double gj1 = (j > 0) ? g[j - 1] : 0.0;
  const int nc = 2 * j + 2;
  if (nc <= kColAThreads / 64) fold_cols<64>(part, nb, nc, val);
  else if (nc <= kColAThreads / 32) fold_cols<32>(part, nb, nc, val);
  else if (nc <= kColAThreads / 16) fold_cols<16>(part, nb, nc, val);
  else fold_cols<8>(part, nb, nc, val);
  if (t < j) {
    rot[t] = rh;
    rot[kMaxFusedGs + t] = rc;
    rot[2 * kMaxFusedGs + t] = rs;
  }
  __syncthreads();
  const double* a = val;
  const double* c = val + j;
  if (t < j) {  // the next provisional column starts from c; a and c are kept for the update pass
    hp[t] = c[t];
    sm[L.h() + t] = a[t];
    sm[L.h2() + t] = c[t];
  }
  if (t != 0) return;
  const double ss = val[2 * j], dd = val[2 * j + 1];
  double aa = 0.0, ac = 0.0;
  for (int q = 0; q < j; ++q) {
    aa += a[q] * a[q];
    ac += a[q] * c[q];
  }
  const double nu = (j == 0) ? 1.0 : sqrt(fmax(ss - aa, 0.0));
  if (j > 0) {  // column j-1 is final: hp + a over rows 0..j-1, nu in row j; rotations 0..j-2, then a new one
    const double* oh = rot;
    const double* oc = rot + kMaxFusedGs;
    const double* os = rot + 2 * kMaxFusedGs;
    double x = oh[0] + a[0];
    for (int q = 0; q + 1 < j; ++q) {
      const double y = oh[q + 1] + a[q + 1];
      sm[L.H(q, j - 1)] = oc[q] * x + os[q] * y;
      x = -os[q] * x + oc[q] * y;
    }
    double cr, sr;
    givens(x, nu, cr, sr);
    cs[j - 1] = cr;
    sn[j - 1] = sr;
    sm[L.H(j - 1, j - 1)] = cr * x + sr * nu;
    sm[L.H(j, j - 1)] = 0.0;
    g[j] = -sr * gj1;
    g[j - 1] = cr * gj1;
  }
  const double e = (nu > 0.0) ? (dd - ac) / nu : 0.0;
  hp[j] = e;
  sm[L.lag()] = (nu > 0.0) ? 1.0 / nu : 0.0;
  sm[L.lag() + 1] = e;
}

// residual estimate of step j from its provisional column (hp, ||w1||): stored in
// sc[2] and in *est (pinned host memory)
__global__ __launch_bounds__(kColAThreads) void k_lag_colB(Layout L, int j, double* __restrict__ sm,
                                                           const double* __restrict__ part, int nb, double* est) {
  __shared__ double lds[kColAThreads / kWave];
  __shared__ double rot[3 * kMaxFusedGs + 1];  // hp rows 0..j, cs and sn rows 0..j-1
  const int t = threadIdx.x;
  const double* hp = sm + L.hp();
  const double* cs = sm + L.cs();
  const double* sn = sm + L.sn();
  double rh = 0.0, rc = 0.0, rs = 0.0;
  if (t <= j) rh = hp[t];
  if (t < j) {
    rc = cs[t];
    rs = sn[t];
  }
  const double gj = sm[L.g() + j];
  double v = 0.0;
  for (int b = t; b < nb; b += blockDim.x) v += part[b];
  if (t <= j) rot[t] = rh;
  if (t < j) {
    rot[kMaxFusedGs + 1 + t] = rc;
    rot[2 * kMaxFusedGs + 1 + t] = rs;
  }
  const double ww = block_sum(v, lds);  // its barrier also publishes rot
  if (t != 0) return;
  const double* oc = rot + kMaxFusedGs + 1;
  const double* os = rot + 2 * kMaxFusedGs + 1;
  double x = rot[0];
  for (int q = 0; q < j; ++q) {  // rotations 0..j-1 applied to the column; only its last two entries matter
    const double y = rot[q + 1];
    x = -os[q] * x + oc[q] * y;
  }
  double cr, sr;
  givens(x, sqrt(fmax(ww, 0.0)), cr, sr);
  double* sc = sm + L.sc();
  sc[2] = fabs(sr * gj);
  *est = sc[2];
}

template <int K>
void launch_lag_pair(int k, double* V, long long n, const double* w, double* sm, Layout L, const double* S, void* bout,
                     int precision, double* part, bool dots_only, const double* u, hipStream_t s) {
  constexpr int P = K >= 16 ? 2 : 1;  // two lanes per row where the accumulators would crowd the registers
  // two rows per lane with 16-byte loads where n is even and the arrays aligned:
  // half the load instructions for the same bytes (7.70 -> 7.60 ms per step at 1M cells)
  auto a16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool two = n % 2 == 0 && a16(V) && a16(w) && (dots_only ? a16(u) : (a16(S) && bout != nullptr));
  if (dots_only) {
    if (two) k_lag_dots<K, P, 2><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, u, w, part);
    else k_lag_dots<K, P, 1><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, u, w, part);
  } else if (precision == 32) {
    if (two)
      k_lag_update<K, P, float, 2><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, w, sm + L.h(), sm + L.h2(), sm + L.lag(),
                                                                    S, (float*)bout, part);
    else
      k_lag_update<K, P, float, 1><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, w, sm + L.h(), sm + L.h2(), sm + L.lag(),
                                                                    S, (float*)bout, part);
  } else {
    if (two)
      k_lag_update<K, P, double, 2><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, w, sm + L.h(), sm + L.h2(),
                                                                     sm + L.lag(), S, (double*)bout, part);
    else
      k_lag_update<K, P, double, 1><<<kGsBlocks, kGsThreads, 0, s>>>(k, V, n, w, sm + L.h(), sm + L.h2(),
                                                                     sm + L.lag(), S, (double*)bout, part);
  }
  IMPES_LAUNCH_CHECK();
}

// the dots pass (dots_only) or the update pass of step k, by the basis size k
void launch_lag(int k, double* V, long long n, const double* w, double* sm, Layout L, const double* S, void* bout,
                int precision, double* part, bool dots_only, const double* u, hipStream_t s) {
  if (k <= 4) launch_lag_pair<4>(k, V, n, w, sm, L, S, bout, precision, part, dots_only, u, s);
  else if (k <= 8) launch_lag_pair<8>(k, V, n, w, sm, L, S, bout, precision, part, dots_only, u, s);
  else if (k <= 16) launch_lag_pair<16>(k, V, n, w, sm, L, S, bout, precision, part, dots_only, u, s);
  else launch_lag_pair<kMaxFusedGs>(k, V, n, w, sm, L, S, bout, precision, part, dots_only, u, s);
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

// y = H^-1 g for the k x k upper-triangular H, one lane per row (k <= 64): the
// triangle is staged in LDS by column with one round of loads, then column i
// eliminates y_i from the rows above it (one shuffle, divide and multiply-add
// per step, no global round trip on the chain).
constexpr int kBackLanes = 64;
__global__ __launch_bounds__(kBackLanes) void k_backsolve_wave(Layout L, int k, double* __restrict__ sm) {
  __shared__ double hs[kBackLanes * kBackLanes];  // hs[i * 64 + j] = H(j, i), j <= i
  const int lane = threadIdx.x;
  for (int c = 0; c < k; ++c)
    if (lane <= c) hs[c * kBackLanes + lane] = sm[L.H(lane, c)];
  double gj = (lane < k) ? sm[L.g() + lane] : 0.0;
  __syncthreads();
  double* y = sm + L.y();
  for (int i = k - 1; i >= 0; --i) {
    const double gi = __shfl(gj, i, kWave);
    const double d = hs[i * kBackLanes + i];
    const double yi = (d != 0.0) ? gi / d : 0.0;
    if (lane == i) y[i] = yi;
    if (lane < i) gj -= hs[i * kBackLanes + lane] * yi;
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
  const char* g = std::getenv("IMPES_FGMRES_CGS2");
  cgs2_ = g && g[0] == '1';
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
  work_.alloc(std::max({(size_t)kRedBlocks * (m + 1), (size_t)kMultiDotBlocks * (m + 1),
                        (size_t)kGsBlocks * (2 * m + 2)}));
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
  double* V = V_.p;
  double* Z = Z_.p;
  // ||D^-1 b|| and the scaled initial residual r0 = D^-1 (b - A x0) in V, their
  // squared norms read back with one synchronisation
  launch_scale_f64(bs_.p, A.dinv, b, n, s);
  launch_dot_f64(bs_.p, bs_.p, n, sc + 5, work_.p, s);
  if (!use_x0) launch_fill_f64(x, 0.0, n, s);
  residual(A, x, b, V, s);
  launch_dot_f64(V, V, n, sc + 4, work_.p, s);
  IMPES_HIP_CHECK(hipMemcpyAsync(host_ + 4, sc + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(host_[5]);
  double beta = std::sqrt(host_[4]);
  st.bnorm = bnorm;
  if (bnorm == 0.0) {
    launch_fill_f64(x, 0.0, n, s);
    st.converged = 1;
    return;
  }
  const double tol = rtol * bnorm;
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
  st.r0_rel = beta / bnorm;
  int its = 0;
  const bool check = etol_ > 0.0 && M != nullptr;
  // residual target of the Krylov cycles: tol, or lower when the error check
  // predicts that tol leaves too large a pressure error
  double target = tol;
  const double ratio = std::max(err_ratio_[0], err_ratio_[1]);
  if (check && ratio > 0.0) target = std::min(tol, kErrTarget * etol_ / ratio * bnorm);
  bool ok = false;
  // two passes over the basis per Arnoldi step (lagged reorthogonalisation) when
  // a preconditioner stages the next V-cycle input and the basis fits the fused
  // kernels; else CGS2's three passes and a normalisation
  const bool lagged = M != nullptr && m <= kMaxFusedGs && !cgs2_;
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
      // lagged orthogonalisation: columns of H final so far (the lagged loop only)
      int final_cols = 0;
      for (int j = 0; j < m; ++j) {
        double* vj = V + (long long)j * n;
        double* zj = Z + (long long)j * n;
        double* est = host_dev_ + 2 + (j & 1);
        if (lagged) {
          // z_j = M u (the V-cycle input S .* u was written by the previous update,
          // or by the cycle's normalisation for j = 0), w = D^-1 A z_j
          precondition_apply(vj, zj);
          launch_lag(j, V, n, w_.p, sm, L, nullptr, nullptr, precision, work_.p, true, vj, s);
          k_lag_colA<<<1, kColAThreads, 0, s>>>(L, j, sm, work_.p, kGsBlocks);
          IMPES_LAUNCH_CHECK();
          final_cols = j;
          launch_lag(j, V, n, w_.p, sm, L, M->in_scale(), M->vcycle_in(precision), precision, work_.p, false,
                     nullptr, s);
          k_lag_colB<<<1, kColAThreads, 0, s>>>(L, j, sm, work_.p, kGsBlocks, est);
          IMPES_LAUNCH_CHECK();
        } else {
        precondition_apply(vj, zj);
        // CGS2: h = V^T w; then w -= V h with h2 = V^T w in one sweep; then w -= V h2
        // with ||w||^2 and the rotations
        const int kj = j + 1;
        if (kj <= kMaxFusedGs) {
          launch_gs_pass<0>(kj, w_.p, V, n, nullptr, work_.p, s);
          k_fold_cols<<<kj, kGsThreads, 0, s>>>(work_.p, kGsBlocks, sm + L.h());
          IMPES_LAUNCH_CHECK();
          launch_gs_pass<1>(kj, w_.p, V, n, sm + L.h(), work_.p, s);
          k_fold_cols<<<kj, kGsThreads, 0, s>>>(work_.p, kGsBlocks, sm + L.h2());
          IMPES_LAUNCH_CHECK();
          launch_gs_pass<2>(kj, w_.p, V, n, sm + L.h2(), work_.p, s);
        } else {
          launch_multidot_f64(V, n, kj, w_.p, n, sm + L.h(), work_.p, s);
          launch_gs_update_f64(w_.p, V, n, kj, sm + L.h(), n, s);
          launch_multidot_f64(V, n, kj, w_.p, n, sm + L.h2(), work_.p, s);
          launch_gs_update_f64(w_.p, V, n, kj, sm + L.h2(), n, s);
          launch_dot_partial_f64(w_.p, w_.p, n, work_.p, kMultiDotBlocks, s);
        }
        k_arnoldi<<<1, kGsThreads, 0, s>>>(L, j, sm, work_.p, kj <= kMaxFusedGs ? kGsBlocks : kMultiDotBlocks, est);
        IMPES_LAUNCH_CHECK();
        normalize(V + (long long)(j + 1) * n, w_.p, sc + 3);
        }
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
      if (lagged && final_cols < k) {
        // close column k-1: dots of u = V[:, k] against v_0..v_{k-1}, then its colA
        const double* uk = V + (long long)k * n;
        launch_lag(k, V, n, uk, sm, L, nullptr, nullptr, precision, work_.p, true, uk, s);
        k_lag_colA<<<1, kColAThreads, 0, s>>>(L, k, sm, work_.p, kGsBlocks);
        IMPES_LAUNCH_CHECK();
      }
      if (k <= kBackLanes) k_backsolve_wave<<<1, kBackLanes, 0, s>>>(L, k, sm);
      else k_backsolve<<<1, 1, 0, s>>>(L, k, sm);
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
