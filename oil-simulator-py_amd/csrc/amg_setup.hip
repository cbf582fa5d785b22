// Ruge-Stueben AMG setup kernels.
//
// Restates, per row and without host round trips, the setup of
// src/solver/classical_amg.py:
//   strength of connection            find_strong_connections   :34-82
//   C/F split                         classical_coarsening      :85-146
//   orphan promotion + interpolation  build_prolongation_rs_full :149-427
//   Galerkin product                  rap_torch / _csr_spmm     :519-657
//   L1-Jacobi relaxation weights      ClassicalAMG.__init__     :885-889
// The unseeded torch.rand tie-break of the C/F split (:125) is replaced by a
// deterministic integer hash of the row index, so the hierarchy is reproducible.
#include "amg_kernels.h"
#include "prims.h"

namespace impes {

namespace {
constexpr int kBlock = 256;
// Operators with rows longer than this on average (the coarse levels, which have
// few rows) run the row-walking setup kernels with a wavefront per row; shorter
// rows use a thread per row. Both forms give identical results.
constexpr long long kWaveMeanRow = 10;
inline bool wave_rows(int n, long long nnz) { return nnz > kWaveMeanRow * (long long)n; }
// The independent-set rounds stop walking a row at its first deciding entry,
// which favours a thread per row up to longer rows.
inline bool wave_rows_mis(int n, long long nnz) { return nnz > 24 * (long long)n; }
// Product expansion: the lanes of a wavefront per row split each right-operand
// row, so they pay off once those rows are long (products per left entry) or
// when there are too few rows (< 2^18) to fill the GPU with a thread per row.
inline bool wave_rows_expand(int n, long long left_nnz, long long products) {
  return products > 8 * left_nnz || n < (1 << 18);
}

__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// lexicographic (deg, hash, index) comparison: true if node a ranks above node b
__device__ __forceinline__ bool ranks_above(int deg_a, int a, int deg_b, int b) {
  if (deg_a != deg_b) return deg_a > deg_b;
  const unsigned ha = hash32((unsigned)a), hb = hash32((unsigned)b);
  if (ha != hb) return ha > hb;
  return a > b;
}

#define ROW_LOOP(i, n) for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)
#define WAVE_ROW_LOOP(i, n)                                                                  \
  for (long long i = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave; i < (n); \
       i += (long long)gridDim.x * blockDim.x / kWave)

// Eight lanes per row: the row sum of |a_ij| folds with shuffles and the first
// diagonal entry (lowest position) wins, as in the one-thread-per-row definition.
constexpr int kDiagLanes = 8;

__global__ void k_diag_relax(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ v, double* __restrict__ diag, double* __restrict__ ir64,
                             float* __restrict__ ir32) {
  const int g = threadIdx.x % kDiagLanes;
  for (long long i = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kDiagLanes; i < n;
       i += (long long)gridDim.x * blockDim.x / kDiagLanes) {
    double d = 0.0, rs = 0.0;
    int dpos = 0x7fffffff;
    const int e1 = rp[i + 1];
    for (int e = rp[i] + g; e < e1; e += kDiagLanes) {
      rs += fabs(v[e]);
      if (ci[e] == (int)i && e < dpos) { dpos = e; d = v[e]; }
    }
    rs = group_sum<kDiagLanes>(rs);
    for (int off = kDiagLanes / 2; off > 0; off >>= 1) {
      const int op = __shfl_xor(dpos, off, kWave);
      const double od = __shfl_xor(d, off, kWave);
      if (op < dpos) { dpos = op; d = od; }
    }
    if (g != 0) continue;
    if (diag) diag[i] = d;
    double den = fmax(fabs(d), 0.3 * rs);
    den = fmax(den, 1e-30);
    const double ir = fmin(1.0 / den, 1e2);
    if (ir64) ir64[i] = ir;
    if (ir32) ir32[i] = (float)ir;
  }
}

// k_diag_relax fused with the refresh of the 4-entry-padded FP32 values (v4,
// rows starting at rp4, padding zeros): the same lanes walk the row once, in the
// same entry order, so the relaxation weights equal k_diag_relax's.
__global__ void k_diag_relax_pad4(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const double* __restrict__ v, const int* __restrict__ rp4, float* __restrict__ v4,
                                  double* __restrict__ ir64, float* __restrict__ ir32) {
  const int g = threadIdx.x % kDiagLanes;
  for (long long i = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kDiagLanes; i < n;
       i += (long long)gridDim.x * blockDim.x / kDiagLanes) {
    double d = 0.0, rs = 0.0;
    int dpos = 0x7fffffff;
    const int e0 = rp[i], len = rp[i + 1] - e0, q0 = rp4[i], len4 = rp4[i + 1] - q0;
    for (int k = g; k < len4; k += kDiagLanes) {
      float out = 0.f;
      if (k < len) {
        const int e = e0 + k;
        const double a = v[e];
        rs += fabs(a);
        if (ci[e] == (int)i && e < dpos) { dpos = e; d = a; }
        out = (float)a;
      }
      v4[q0 + k] = out;
    }
    rs = group_sum<kDiagLanes>(rs);
    for (int off = kDiagLanes / 2; off > 0; off >>= 1) {
      const int op = __shfl_xor(dpos, off, kWave);
      const double od = __shfl_xor(d, off, kWave);
      if (op < dpos) { dpos = op; d = od; }
    }
    if (g != 0) continue;
    double den = fmax(fabs(d), 0.3 * rs);
    den = fmax(den, 1e-30);
    const double ir = fmin(1.0 / den, 1e2);
    ir64[i] = ir;
    ir32[i] = (float)ir;
  }
}

__global__ void k_small_diag(int n, const double* __restrict__ diag, double thr, int* __restrict__ flags) {
  ROW_LOOP(i, n) flags[i] = (fmax(fabs(diag[i]), 1e-30) < thr) ? 1 : 0;
}

__global__ void k_dhalf_inv(int n, const double* __restrict__ diag, double* __restrict__ dh) {
  ROW_LOOP(i, n) dh[i] = 1.0 / sqrt(fmax(fabs(diag[i]), 1e-30));
}

__global__ void k_sym_scale(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ v, const double* __restrict__ dh, double* __restrict__ vo) {
  ROW_LOOP(i, n) {
    const double di = dh[i];
    for (int e = rp[i]; e < rp[i + 1]; ++e) vo[e] = v[e] * di * dh[ci[e]];
  }
}

__global__ void k_strength(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                           double theta, unsigned char* __restrict__ strong) {
  ROW_LOOP(i, n) {
    double d = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
      if (ci[e] == i) { d = fabs(v[e]); break; }
    const double den = d + 1e-30;
    double mx = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const double sv = v[e] / den;
      if (ci[e] != i && sv < 0.0) mx = fmax(mx, -sv);
    }
    const double thr = theta * mx;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const double sv = v[e] / den;
      strong[e] = (ci[e] != i && sv < 0.0 && -sv >= thr) ? 1 : 0;
    }
  }
}

__global__ void k_transpose_pos(int n, const int* __restrict__ rp, const int* __restrict__ ci, int* __restrict__ tpos) {
  ROW_LOOP(i, n) {
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int j = ci[e];
      int lo = rp[j], hi = rp[j + 1] - 1, pos = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = ci[mid];
        if (c == (int)i) { pos = mid; break; }
        if (c < (int)i) lo = mid + 1; else hi = mid - 1;
      }
      tpos[e] = pos;
    }
  }
}

// Wavefront-per-row forms of k_strength and k_transpose_pos. The diagonal is the
// row's first diagonal entry (lowest position over the lanes) and the largest
// off-diagonal magnitude is a max, so the marks equal the one-thread ones.
__global__ void k_strength_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ v, double theta, unsigned char* __restrict__ strong) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    const int r0 = rp[i], r1 = rp[i + 1];
    int dpos = 0x7fffffff;
    for (int e = r0 + lane; e < r1; e += kWave)
      if (ci[e] == (int)i) { dpos = e; break; }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const int o = __shfl_xor(dpos, off, kWave);
      dpos = o < dpos ? o : dpos;
    }
    const double d = (dpos < r1) ? fabs(v[dpos]) : 0.0;
    const double den = d + 1e-30;
    double mx = 0.0;
    for (int e = r0 + lane; e < r1; e += kWave) {
      const double sv = v[e] / den;
      if (ci[e] != (int)i && sv < 0.0) mx = fmax(mx, -sv);
    }
    mx = wave_max(mx);
    const double thr = theta * mx;
    for (int e = r0 + lane; e < r1; e += kWave) {
      const double sv = v[e] / den;
      strong[e] = (ci[e] != (int)i && sv < 0.0 && -sv >= thr) ? 1 : 0;
    }
  }
}

__global__ void k_transpose_pos_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                     int* __restrict__ tpos) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    for (int e = rp[i] + lane; e < rp[i + 1]; e += kWave) {
      const int j = ci[e];
      int lo = rp[j], hi = rp[j + 1] - 1, pos = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = ci[mid];
        if (c == (int)i) { pos = mid; break; }
        if (c < (int)i) lo = mid + 1; else hi = mid - 1;
      }
      tpos[e] = pos;
    }
  }
}

__device__ __forceinline__ bool in_graph(const unsigned char* strong, const int* tpos, int e) {
  const int t = tpos[e];
  return strong[e] || (t >= 0 && strong[t]);
}

__global__ void k_cf_deg(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                         const unsigned char* __restrict__ strong, const int* __restrict__ tpos, int* __restrict__ deg) {
  ROW_LOOP(i, n) {
    int d = 0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
      if (ci[e] != (int)i && in_graph(strong, tpos, e)) ++d;
    deg[i] = d;
  }
}

__global__ void k_cf_select(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                            const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                            const int* __restrict__ deg, signed char* __restrict__ is_max) {
  ROW_LOOP(i, n) {
    const int di = deg[i];
    bool mx = true;
    for (int e = rp[i]; e < rp[i + 1] && mx; ++e) {
      const int j = ci[e];
      if (j != (int)i && in_graph(strong, tpos, e) && ranks_above(deg[j], j, di, (int)i)) mx = false;
    }
    is_max[i] = mx ? 1 : 0;
  }
}

__global__ void k_cf_lift(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                          const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                          const signed char* __restrict__ is_max, signed char* __restrict__ cf) {
  ROW_LOOP(i, n) {
    if (is_max[i]) { cf[i] = kC; continue; }
    bool has_c = false;
    for (int e = rp[i]; e < rp[i + 1] && !has_c; ++e) {
      const int j = ci[e];
      if (j != (int)i && in_graph(strong, tpos, e) && is_max[j]) has_c = true;
    }
    cf[i] = has_c ? kF : kC;
  }
}

// Iterated independent-set split (not in the reference): every round selects the
// undecided points that rank above all their undecided graph neighbours as C, and
// turns the undecided neighbours of a new C point into F, until none is undecided.
constexpr signed char kUndecided = 2;

__global__ void k_mis_init(int n, signed char* __restrict__ cf) {
  ROW_LOOP(i, n) cf[i] = kUndecided;
}

__global__ void k_mis_select(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                             const int* __restrict__ deg, const signed char* __restrict__ cf,
                             signed char* __restrict__ newc) {
  ROW_LOOP(i, n) {
    bool mx = cf[i] == kUndecided;
    const int di = deg[i];
    for (int e = rp[i]; e < rp[i + 1] && mx; ++e) {
      const int j = ci[e];
      if (j != (int)i && cf[j] == kUndecided && in_graph(strong, tpos, e) && ranks_above(deg[j], j, di, (int)i))
        mx = false;
    }
    newc[i] = mx ? 1 : 0;
  }
}

__global__ void k_mis_cover(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                            const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                            const signed char* __restrict__ newc, signed char* __restrict__ cf,
                            int* __restrict__ undecided) {
  ROW_LOOP(i, n) {
    int u = 0;
    if (newc[i]) {
      cf[i] = kC;
    } else if (cf[i] == kUndecided) {
      bool cov = false;
      for (int e = rp[i]; e < rp[i + 1] && !cov; ++e) {
        const int j = ci[e];
        if (j != (int)i && newc[j] && in_graph(strong, tpos, e)) cov = true;
      }
      if (cov) cf[i] = kF;
      else u = 1;
    }
    undecided[i] = u;
  }
}

// Wavefront-per-row forms of k_cf_deg, k_mis_select and k_mis_cover: counts
// summed over the lanes, existence tests by ballot; the same results.
__global__ void k_cf_deg_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                              int* __restrict__ deg) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    int d = 0;
    for (int e = rp[i] + lane; e < rp[i + 1]; e += kWave)
      if (ci[e] != (int)i && in_graph(strong, tpos, e)) ++d;
    d = wave_sum(d);
    if (lane == 0) deg[i] = d;
  }
}

__global__ void k_mis_select_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                                  const int* __restrict__ deg, const signed char* __restrict__ cf,
                                  signed char* __restrict__ newc) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    bool beaten = false;
    if (cf[i] == kUndecided) {
      const int di = deg[i];
      for (int e = rp[i] + lane; e < rp[i + 1] && !beaten; e += kWave) {
        const int j = ci[e];
        if (j != (int)i && cf[j] == kUndecided && in_graph(strong, tpos, e) && ranks_above(deg[j], j, di, (int)i))
          beaten = true;
      }
      beaten = __ballot(beaten) != 0;
      if (lane == 0) newc[i] = beaten ? 0 : 1;
    } else if (lane == 0) {
      newc[i] = 0;
    }
  }
}

__global__ void k_mis_cover_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const unsigned char* __restrict__ strong, const int* __restrict__ tpos,
                                 const signed char* __restrict__ newc, signed char* __restrict__ cf,
                                 int* __restrict__ undecided) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    int u = 0;
    const signed char ci0 = cf[i];
    if (newc[i]) {
      if (lane == 0) cf[i] = kC;
    } else if (ci0 == kUndecided) {
      bool cov = false;
      for (int e = rp[i] + lane; e < rp[i + 1] && !cov; e += kWave) {
        const int j = ci[e];
        if (j != (int)i && newc[j] && in_graph(strong, tpos, e)) cov = true;
      }
      cov = __ballot(cov) != 0;
      if (lane == 0) {
        if (cov) cf[i] = kF;
        else u = 1;
      }
    }
    if (lane == 0) undecided[i] = u;
  }
}

__global__ void k_force_dominant(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const double* __restrict__ v, double ratio, signed char* __restrict__ cf) {
  ROW_LOOP(i, n) {
    double d = 0.0, off = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      if (ci[e] == (int)i) d += v[e];
      else off += fabs(v[e]);
    }
    if (fabs(d) > ratio * off) cf[i] = kC;
  }
}

__global__ void k_orphans(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                          const unsigned char* __restrict__ strong, const signed char* __restrict__ cf_in,
                          signed char* __restrict__ cf_out, int* __restrict__ is_c) {
  ROW_LOOP(i, n) {
    signed char c = cf_in[i];
    if (c == kF) {
      bool has = false;
      for (int e = rp[i]; e < rp[i + 1] && !has; ++e) {
        const int j = ci[e];
        if (j != (int)i && strong[e] && cf_in[j] == kC) has = true;
      }
      if (!has) c = kC;
    }
    cf_out[i] = c;
    is_c[i] = (c == kC) ? 1 : 0;
  }
}

// ---------------------------------------------------------------- interpolation
__global__ void k_interp_a(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                           const unsigned char* __restrict__ strong, const signed char* __restrict__ cf,
                           double* __restrict__ sum_strong, int* __restrict__ cnt_c) {
  ROW_LOOP(i, n) {
    double s = 0.0;
    int c = 0;
    if (cf[i] == kF) {
      for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (j != (int)i && strong[e] && cf[j] == kC) { s += -v[e]; ++c; }
      }
    }
    sum_strong[i] = s;
    cnt_c[i] = c;
  }
}

__device__ __forceinline__ bool ff_valid(const double* sum_strong, const int* cnt_c, int j) {
  return cnt_c[j] > 0 && sum_strong[j] > 1e-14;
}

__global__ void k_interp_count(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const unsigned char* __restrict__ strong, const signed char* __restrict__ cf,
                               const double* __restrict__ sum_strong, const int* __restrict__ cnt_c,
                               int* __restrict__ cand) {
  ROW_LOOP(i, n) {
    if (cf[i] == kC) { cand[i] = 1; continue; }
    int c = cnt_c[i];
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int j = ci[e];
      if (j != (int)i && strong[e] && cf[j] == kF && ff_valid(sum_strong, cnt_c, j)) c += cnt_c[j];
    }
    // one spare slot (sentinel key) for the zero-row fallback insertion
    cand[i] = c + 1;
  }
}

__device__ __forceinline__ double finite_or_zero(double x) { return isfinite(x) ? x : 0.0; }

__global__ void k_interp_expand(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ v, const unsigned char* __restrict__ strong,
                                const signed char* __restrict__ cf, const double* __restrict__ sum_strong,
                                const int* __restrict__ cnt_c, const int* __restrict__ f2c,
                                const long long* __restrict__ cand_ptr, unsigned long long* __restrict__ keys,
                                double* __restrict__ vals) {
  ROW_LOOP(i, n) {
    const long long base = cand_ptr[i];
    if (cf[i] == kC) {
      keys[base] = ((unsigned long long)(unsigned)f2c[i]) << 32;
      vals[base] = 1.0;
      continue;
    }
    double diag = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
      if (ci[e] == (int)i) { diag += v[e]; }
    if (!(fabs(diag) > 1e-14)) diag = 1.0;
    double weak_c = 0.0, f_weak = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int j = ci[e];
      if (j == (int)i) continue;
      if (!strong[e]) {
        if (cf[j] == kC) weak_c += -v[e];
        else f_weak += -v[e];
      } else if (cf[j] == kF && !ff_valid(sum_strong, cnt_c, j)) {
        f_weak += -v[e];
      }
    }
    const double ss = sum_strong[i];
    const double corr = (ss > 1e-14) ? (weak_c + f_weak) / ss : 0.0;
    unsigned local = 0;
    // direct strong C neighbours
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int j = ci[e];
      if (j != (int)i && strong[e] && cf[j] == kC) {
        double w = finite_or_zero(-v[e] / diag) * (1.0 + corr);
        w = fmax(w, 0.0);
        keys[base + local] = (((unsigned long long)(unsigned)f2c[j]) << 32) | local;
        vals[base + local] = w;
        ++local;
      }
    }
    // distance-two contributions through strong F neighbours
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int j = ci[e];
      if (j == (int)i || !strong[e] || cf[j] != kF || !ff_valid(sum_strong, cnt_c, j)) continue;
      const double factor = finite_or_zero(-v[e] / diag);
      const double ssj = sum_strong[j];
      for (int e2 = rp[j]; e2 < rp[j + 1]; ++e2) {
        const int k = ci[e2];
        if (k != j && strong[e2] && cf[k] == kC) {
          const double jw = -v[e2] / ssj;
          const double c = fmax(factor * jw, 0.0);
          keys[base + local] = (((unsigned long long)(unsigned)f2c[k]) << 32) | local;
          vals[base + local] = c;
          ++local;
        }
      }
    }
    keys[base + local] = ~0ull;  // spare slot, sorts last
    vals[base + local] = 0.0;
  }
}

// Lane q's value of x (q uniform over the wavefront).
__device__ __forceinline__ double read_lane(double x, int q) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, q);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), q);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int wave_inclusive_scan(int x, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int o = __shfl_up(x, off, kWave);
    if (lane >= off) x += o;
  }
  return x;
}

// Wavefront-per-row form of k_interp_expand for long rows (coarse levels),
// with the same candidates, weights and slots. The lanes load the row's entries
// side by side; the diagonal and weak-connection sums fold lane by lane in row
// order, so they equal the one-thread sums bit for bit. Direct C slots come
// from ballot ranks; each valid strong F neighbour gets a block of cnt_c slots
// from a wavefront prefix sum, which its lane fills from that neighbour's row.
__global__ void k_interp_expand_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                     const double* __restrict__ v, const unsigned char* __restrict__ strong,
                                     const signed char* __restrict__ cf, const double* __restrict__ sum_strong,
                                     const int* __restrict__ cnt_c, const int* __restrict__ f2c,
                                     const long long* __restrict__ cand_ptr, unsigned long long* __restrict__ keys,
                                     double* __restrict__ vals) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  WAVE_ROW_LOOP(i, n) {
    const long long base = cand_ptr[i];
    if (cf[i] == kC) {
      if (lane == 0) {
        keys[base] = ((unsigned long long)(unsigned)f2c[i]) << 32;
        vals[base] = 1.0;
      }
      continue;
    }
    const int r0 = rp[i], r1 = rp[i + 1];
    double diag = 0.0, weak_c = 0.0, f_weak = 0.0;
    for (int e0 = r0; e0 < r1; e0 += kWave) {
      const int e = e0 + lane;
      int cat = 0;  // 1 diagonal, 2 weak C, 3 weak F or strong F without C support
      double x = 0.0;
      if (e < r1) {
        const int j = ci[e];
        if (j == (int)i) {
          cat = 1;
          x = v[e];
        } else if (!strong[e]) {
          cat = (cf[j] == kC) ? 2 : 3;
          x = -v[e];
        } else if (cf[j] == kF && !ff_valid(sum_strong, cnt_c, j)) {
          cat = 3;
          x = -v[e];
        }
      }
      const int m = min(kWave, r1 - e0);
      for (int q = 0; q < m; ++q) {
        const int c = __builtin_amdgcn_readlane(cat, q);
        if (c == 0) continue;
        const double y = read_lane(x, q);
        if (c == 1) diag += y;
        else if (c == 2) weak_c += y;
        else f_weak += y;
      }
    }
    if (!(fabs(diag) > 1e-14)) diag = 1.0;
    const double ss = sum_strong[i];
    const double corr = (ss > 1e-14) ? (weak_c + f_weak) / ss : 0.0;
    unsigned local = 0;
    for (int e0 = r0; e0 < r1; e0 += kWave) {
      const int e = e0 + lane;
      int j = 0;
      bool take = false;
      if (e < r1) {
        j = ci[e];
        take = j != (int)i && strong[e] && cf[j] == kC;
      }
      const unsigned long long m = __ballot(take);
      if (take) {
        const unsigned slot = local + (unsigned)__popcll(m & below);
        double w = finite_or_zero(-v[e] / diag) * (1.0 + corr);
        w = fmax(w, 0.0);
        keys[base + slot] = (((unsigned long long)(unsigned)f2c[j]) << 32) | slot;
        vals[base + slot] = w;
      }
      local += (unsigned)__popcll(m);
    }
    for (int e0 = r0; e0 < r1; e0 += kWave) {
      const int e = e0 + lane;
      int j = 0, cnt = 0;
      if (e < r1) {
        j = ci[e];
        if (j != (int)i && strong[e] && cf[j] == kF && ff_valid(sum_strong, cnt_c, j)) cnt = cnt_c[j];
      }
      const int incl = wave_inclusive_scan(cnt, lane);
      if (cnt > 0) {
        unsigned slot = local + (unsigned)(incl - cnt);
        const double factor = finite_or_zero(-v[e] / diag);
        const double ssj = sum_strong[j];
        for (int e2 = rp[j]; e2 < rp[j + 1]; ++e2) {
          const int k = ci[e2];
          if (k != j && strong[e2] && cf[k] == kC) {
            const double jw = -v[e2] / ssj;
            const double c = fmax(factor * jw, 0.0);
            keys[base + slot] = (((unsigned long long)(unsigned)f2c[k]) << 32) | slot;
            vals[base + slot] = c;
            ++slot;
          }
        }
      }
      local += (unsigned)__shfl(incl, kWave - 1, kWave);
    }
    if (lane == 0) {
      keys[base + local] = ~0ull;  // spare slot, sorts last
      vals[base + local] = 0.0;
    }
  }
}

__global__ void k_interp_reduce(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const signed char* __restrict__ cf, const int* __restrict__ f2c,
                                const long long* __restrict__ cand_ptr, const unsigned long long* __restrict__ keys,
                                const double* __restrict__ vals, double trunc, int pmax, int* __restrict__ oc,
                                double* __restrict__ ov, int* __restrict__ pcnt) {
  ROW_LOOP(i, n) {
    const long long b = cand_ptr[i], e_end = cand_ptr[i + 1];
    if (cf[i] == kC) {
      oc[b] = f2c[i];
      ov[b] = 1.0;
      pcnt[i] = 1;
      continue;
    }
    // merge duplicates (segment sorted by coarse column, ties in generation order)
    int u = 0;
    double row_sum = 0.0;
    for (long long t = b; t < e_end && keys[t] != ~0ull;) {
      const int col = (int)(keys[t] >> 32);
      double s = 0.0;
      while (t < e_end && keys[t] != ~0ull && (int)(keys[t] >> 32) == col) {
        s += vals[b + (long long)(keys[t] & 0xffffffffull)];
        ++t;
      }
      oc[b + u] = col;
      ov[b + u] = s;
      row_sum += s;
      ++u;
    }
    double scale = 1.0;
    if (row_sum > 1e-14) {
      scale = 1.0 / row_sum;
    } else {
      // no usable weight: inject to the first coarse neighbour of the row, keeping
      // whatever (tiny) weights exist, then renormalise (classical_amg.py:341-384)
      int target = -1;
      for (int e = rp[i]; e < rp[i + 1]; ++e)
        if (cf[ci[e]] == kC) { target = f2c[ci[e]]; break; }
      if (target >= 0) {
        // insert (target, +1) keeping the row sorted
        int pos = 0;
        while (pos < u && oc[b + pos] < target) ++pos;
        if (pos < u && oc[b + pos] == target) {
          ov[b + pos] += 1.0;
        } else {
          // the segment of an F row has one spare slot beyond its candidates, so
          // u + 1 <= segment length here
          for (int q = u; q > pos; --q) { oc[b + q] = oc[b + q - 1]; ov[b + q] = ov[b + q - 1]; }
          oc[b + pos] = target;
          ov[b + pos] = 1.0;
          ++u;
        }
        double rs = 0.0;
        for (int q = 0; q < u; ++q) rs += ov[b + q];
        scale = (rs > 1e-14) ? 1.0 / rs : 1.0;
      }
    }
    int w = 0;
    for (int q = 0; q < u; ++q) {
      const double val = fmax(ov[b + q] * scale, 0.0);
      if (val != 0.0) {
        oc[b + w] = oc[b + q];
        ov[b + w] = val;
        ++w;
      }
    }
    if (w > 1 && (trunc > 0.0 || pmax > 0)) {
      // truncation (not in the reference). Weights within tie = 1e-10 * row max of each
      // other count as equal (ties go to the lower coarse column), so that rounding
      // differences of a few ulps cannot change which weights survive.
      double tot = 0.0, mx = 0.0;
      for (int q = 0; q < w; ++q) {
        tot += ov[b + q];
        mx = fmax(mx, ov[b + q]);
      }
      const double tie = 1e-10 * mx;
      const double cut = trunc * mx;
      int k = 0;
      for (int q = 0; q < w; ++q)
        if (ov[b + q] + tie >= cut) {
          oc[b + k] = oc[b + q];
          ov[b + k] = ov[b + q];
          ++k;
        }
      if (pmax > 0 && k > pmax) {
        // rank of q = entries above it (by more than tie) + tied entries to its left;
        // ranks only read values, so dropped entries are flagged in the column array
        for (int q = 0; q < k; ++q) {
          const double vq = ov[b + q];
          int rank = 0;
          for (int r = 0; r < k; ++r) {
            const double vr = ov[b + r];
            rank += (vr > vq + tie || (fabs(vr - vq) <= tie && r < q)) ? 1 : 0;
          }
          if (rank >= pmax) oc[b + q] = -1 - oc[b + q];
        }
        int m = 0;
        for (int q = 0; q < k; ++q)
          if (oc[b + q] >= 0) {
            oc[b + m] = oc[b + q];
            ov[b + m] = ov[b + q];
            ++m;
          }
        k = m;
      }
      double ks = 0.0;
      for (int q = 0; q < k; ++q) ks += ov[b + q];
      const double f = tot / ks;
      for (int q = 0; q < k; ++q) ov[b + q] = ov[b + q] * f;
      w = k;
    }
    pcnt[i] = w;
  }
}

__global__ void k_interp_gather(int n, const long long* __restrict__ cand_ptr, const int* __restrict__ oc,
                                const double* __restrict__ ov, const int* __restrict__ prp, int* __restrict__ pci,
                                double* __restrict__ pv) {
  ROW_LOOP(i, n) {
    const long long b = cand_ptr[i];
    const int o = prp[i], cnt = prp[i + 1] - prp[i];
    for (int q = 0; q < cnt; ++q) {
      pci[o + q] = oc[b + q];
      pv[o + q] = ov[b + q];
    }
  }
}

// ---------------------------------------------------------------- transpose
__global__ void k_col_count(int nnz, const int* __restrict__ ci, int* __restrict__ cnt) {
  ROW_LOOP(e, nnz) atomicAdd(&cnt[ci[e]], 1);
}

__global__ void k_transpose_fill(int n_rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const int* __restrict__ rrp, int* __restrict__ cursor,
                                 unsigned long long* __restrict__ keys) {
  ROW_LOOP(i, n_rows) {
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int c = ci[e];
      const int slot = atomicAdd(&cursor[c], 1);
      keys[rrp[c] + slot] = (((unsigned long long)(unsigned)i) << 32) | (unsigned)e;
    }
  }
}

__global__ void k_transpose_finish(int nnz, const unsigned long long* __restrict__ keys,
                                   const double* __restrict__ src_v, int* __restrict__ rci, double* __restrict__ rv) {
  ROW_LOOP(p, nnz) {
    const unsigned long long k = keys[p];
    rci[p] = (int)(k >> 32);
    rv[p] = src_v[k & 0xffffffffull];
  }
}

// ---------------------------------------------------------------- SpGEMM
__global__ void k_spgemm_count(int n, const int* __restrict__ lrp, const int* __restrict__ lci,
                               const int* __restrict__ mrp, int* __restrict__ cnt) {
  ROW_LOOP(i, n) {
    int c = 0;
    for (int e = lrp[i]; e < lrp[i + 1]; ++e) {
      const int k = lci[e];
      c += mrp[k + 1] - mrp[k];
    }
    cnt[i] = c;
  }
}

// One wavefront per output row: lanes cover the right-operand row of each left
// entry; every product records its key (column, local index) and its operand
// pair (left entry, right entry) in expansion order.

// The same expansion with a thread per output row (rows with few products).
__global__ void k_spgemm_expand_rows(int n, const int* __restrict__ lrp, const int* __restrict__ lci,
                                     const int* __restrict__ mrp, const int* __restrict__ mci,
                                     const long long* __restrict__ prod_ptr, unsigned long long* __restrict__ keys,
                                     int* __restrict__ ex_e, int* __restrict__ ex_f) {
  ROW_LOOP(i, n) {
    long long t = prod_ptr[i];
    unsigned local = 0;
    for (int e = lrp[i]; e < lrp[i + 1]; ++e) {
      const int k = lci[e];
      const int f0 = mrp[k], len = mrp[k + 1] - f0;
      for (int q = 0; q < len; ++q) {
        keys[t + q] = (((unsigned long long)(unsigned)mci[f0 + q]) << 32) | (unsigned long long)(local + q);
        ex_e[t + q] = e;
        ex_f[t + q] = f0 + q;
      }
      t += len;
      local += (unsigned)len;
    }
  }
}

__global__ void k_spgemm_expand(int n, const int* __restrict__ lrp, const int* __restrict__ lci,
                                const int* __restrict__ mrp, const int* __restrict__ mci,
                                const long long* __restrict__ prod_ptr, unsigned long long* __restrict__ keys,
                                int* __restrict__ ex_e, int* __restrict__ ex_f) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    long long t = prod_ptr[i];
    unsigned local = 0;
    for (int e = lrp[i]; e < lrp[i + 1]; ++e) {
      const int k = lci[e];
      const int f0 = mrp[k], len = mrp[k + 1] - f0;
      for (int q = lane; q < len; q += kWave) {
        keys[t + q] = (((unsigned long long)(unsigned)mci[f0 + q]) << 32) | (unsigned long long)(local + q);
        ex_e[t + q] = e;
        ex_f[t + q] = f0 + q;
      }
      t += len;
      local += (unsigned)len;
    }
  }
}

// distinct columns per row of the sorted product list (a wavefront per row)
__global__ void k_spgemm_unique(int n, const long long* __restrict__ prod_ptr,
                                const unsigned long long* __restrict__ keys, int* __restrict__ cnt) {
  const int lane = threadIdx.x & (kWave - 1);
  WAVE_ROW_LOOP(i, n) {
    const long long b = prod_ptr[i], e_end = prod_ptr[i + 1];
    int c = 0;
    for (long long t0 = b; t0 < e_end; t0 += kWave) {
      const long long t = t0 + lane;
      const bool valid = t < e_end;
      const int col = valid ? (int)(keys[t] >> 32) : 0;
      const int prev = (valid && t > b) ? (int)(keys[t - 1] >> 32) : -1;
      c += __popcll(__ballot(valid && col != prev));
    }
    if (lane == 0) cnt[i] = c;
  }
}

// Output columns and product segments (wave per row: a ballot marks the first
// product of each column, its rank gives the output slot), and the operand pair
// of every sorted product, gathered from the expansion-order record.
__global__ void k_spgemm_finalize(int n, const long long* __restrict__ prod_ptr,
                                  const unsigned long long* __restrict__ keys, const int* __restrict__ crp,
                                  const int* __restrict__ ex_e, const int* __restrict__ ex_f, int* __restrict__ cci,
                                  long long* __restrict__ seg, int* __restrict__ pe, int* __restrict__ pf,
                                  long long total, int nnz_c) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  WAVE_ROW_LOOP(i, n) {
    const long long b = prod_ptr[i], e_end = prod_ptr[i + 1];
    int o = crp[i];
    for (long long t0 = b; t0 < e_end; t0 += kWave) {
      const long long t = t0 + lane;
      const bool valid = t < e_end;
      const unsigned long long key = valid ? keys[t] : 0ull;
      const int col = (int)(key >> 32);
      const int prev = (valid && t > b) ? (int)(keys[t - 1] >> 32) : -1;
      const bool start = valid && col != prev;
      const unsigned long long m = __ballot(start);
      if (valid) {
        const long long src = b + (long long)(key & 0xffffffffull);
        pe[t] = ex_e[src];
        pf[t] = ex_f[src];
      }
      if (start) {
        const int pos = o + __popcll(m & below);
        cci[pos] = col;
        seg[pos] = t;
      }
      o += __popcll(m);
    }
    if (i == n - 1 && lane == 0) seg[nnz_c] = total;
  }
}

// Numeric Galerkin product: output entry o is the sum of lv[pe[t]] * mv[pf[t]]
// over its product terms t in [seg[o], seg[o+1]). G lanes share an output
// entry (G from the mean number of terms, as for the CSR sweeps): lane g takes
// terms g, g+G, ... and the partial sums fold across the G lanes, in a fixed
// order.
template <int G>
__global__ __launch_bounds__(kBlock) void k_spgemm_numeric(int nnz_c, const long long* __restrict__ seg,
                                                           const int* __restrict__ pe, const int* __restrict__ pf,
                                                           const double* __restrict__ lv,
                                                           const double* __restrict__ mv, double* __restrict__ cv) {
  constexpr int OPB = kBlock / G;
  const int g = threadIdx.x % G;
  for (long long base = blockIdx.x * (long long)OPB; base < nnz_c; base += (long long)gridDim.x * OPB) {
    const long long o = base + threadIdx.x / G;
    double s = 0.0;
    if (o < nnz_c) {
      const long long t1 = seg[o + 1];
#pragma unroll 4
      for (long long t = seg[o] + g; t < t1; t += G) s += lv[pe[t]] * mv[pf[t]];
    }
    if (G > 1) s = group_sum<G>(s);
    if (o < nnz_c && g == 0) cv[o] = s;
  }
}

template <class K, class... A>
void launch_rows(K kernel, long long n, hipStream_t s, A... args) {
  if (n <= 0) return;
  kernel<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(args...);
  IMPES_LAUNCH_CHECK();
}

// one wavefront per row
template <class K, class... A>
void launch_waves(K kernel, long long n, hipStream_t s, A... args) {
  if (n <= 0) return;
  kernel<<<grid_for(n * kWave, kBlock, 1 << 16), kBlock, 0, s>>>(args...);
  IMPES_LAUNCH_CHECK();
}

}  // namespace

void launch_diag_relax(int n, const int* rp, const int* ci, const double* v, double* diag, double* ir64, float* ir32,
                       hipStream_t s) {
  if (n <= 0) return;
  k_diag_relax<<<grid_for((long long)n * kDiagLanes, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp, ci, v, diag, ir64, ir32);
  IMPES_LAUNCH_CHECK();
}
void launch_diag_relax_pad4(int n, const int* rp, const int* ci, const double* v, const int* rp4, float* v4,
                            double* ir64, float* ir32, hipStream_t s) {
  if (n <= 0) return;
  k_diag_relax_pad4<<<grid_for((long long)n * kDiagLanes, kBlock, 1 << 16), kBlock, 0, s>>>(n, rp, ci, v, rp4, v4,
                                                                                          ir64, ir32);
  IMPES_LAUNCH_CHECK();
}
void launch_count_small_diag(int n, const double* diag, double thr, int* flags, hipStream_t s) {
  launch_rows(k_small_diag, n, s, n, diag, thr, flags);
}
void launch_dhalf_inv(int n, const double* diag, double* dh, hipStream_t s) { launch_rows(k_dhalf_inv, n, s, n, diag, dh); }
void launch_sym_scale(int n, const int* rp, const int* ci, const double* v, const double* dh, double* vo,
                      hipStream_t s) {
  launch_rows(k_sym_scale, n, s, n, rp, ci, v, dh, vo);
}
void launch_strength(int n, long long nnz, const int* rp, const int* ci, const double* v, double theta,
                     unsigned char* strong, hipStream_t s) {
  if (wave_rows(n, nnz)) launch_waves(k_strength_wave, n, s, n, rp, ci, v, theta, strong);
  else launch_rows(k_strength, n, s, n, rp, ci, v, theta, strong);
}
void launch_transpose_pos(int n, long long nnz, const int* rp, const int* ci, int* tpos, hipStream_t s) {
  if (wave_rows(n, nnz)) launch_waves(k_transpose_pos_wave, n, s, n, rp, ci, tpos);
  else launch_rows(k_transpose_pos, n, s, n, rp, ci, tpos);
}
void launch_cf_split(int n, long long nnz, const int* rp, const int* ci, const unsigned char* strong,
                     const int* tpos, int* deg, signed char* is_max, signed char* cf, hipStream_t s) {
  if (wave_rows(n, nnz)) launch_waves(k_cf_deg_wave, n, s, n, rp, ci, strong, tpos, deg);
  else launch_rows(k_cf_deg, n, s, n, rp, ci, strong, tpos, deg);
  launch_rows(k_cf_select, n, s, n, rp, ci, strong, tpos, (const int*)deg, is_max);
  launch_rows(k_cf_lift, n, s, n, rp, ci, strong, tpos, (const signed char*)is_max, cf);
}
void launch_cf_split_iterated(int n, long long nnz, const int* rp, const int* ci, const unsigned char* strong,
                              const int* tpos, int* deg, signed char* newc, int* undecided, signed char* cf,
                              hipStream_t s) {
  const bool wave = wave_rows_mis(n, nnz);
  if (wave) launch_waves(k_cf_deg_wave, n, s, n, rp, ci, strong, tpos, deg);
  else launch_rows(k_cf_deg, n, s, n, rp, ci, strong, tpos, deg);
  launch_rows(k_mis_init, n, s, n, cf);
  // the undecided point of highest rank is selected in every round, so at most n rounds
  for (long long round = 0;; ++round) {
    IMPES_REQUIRE(round <= n, "amg setup: independent-set split does not terminate");
    if (wave) {
      launch_waves(k_mis_select_wave, n, s, n, rp, ci, strong, tpos, (const int*)deg, (const signed char*)cf, newc);
      launch_waves(k_mis_cover_wave, n, s, n, rp, ci, strong, tpos, (const signed char*)newc, cf, undecided);
    } else {
      launch_rows(k_mis_select, n, s, n, rp, ci, strong, tpos, (const int*)deg, (const signed char*)cf, newc);
      launch_rows(k_mis_cover, n, s, n, rp, ci, strong, tpos, (const signed char*)newc, cf, undecided);
    }
    if (sum_i32(undecided, n, s) == 0) break;
  }
}
void launch_force_dominant(int n, const int* rp, const int* ci, const double* v, double ratio, signed char* cf,
                           hipStream_t s) {
  launch_rows(k_force_dominant, n, s, n, rp, ci, v, ratio, cf);
}
void launch_orphans(int n, const int* rp, const int* ci, const unsigned char* strong, const signed char* cf_in,
                    signed char* cf_out, int* is_c, hipStream_t s) {
  launch_rows(k_orphans, n, s, n, rp, ci, strong, cf_in, cf_out, is_c);
}
void launch_interp_pass_a(int n, const int* rp, const int* ci, const double* v, const unsigned char* strong,
                          const signed char* cf, double* sum_strong, int* cnt_c, hipStream_t s) {
  launch_rows(k_interp_a, n, s, n, rp, ci, v, strong, cf, sum_strong, cnt_c);
}
void launch_interp_count(int n, const int* rp, const int* ci, const unsigned char* strong, const signed char* cf,
                         const double* sum_strong, const int* cnt_c, int* cand, hipStream_t s) {
  launch_rows(k_interp_count, n, s, n, rp, ci, strong, cf, sum_strong, cnt_c, cand);
}
void launch_interp_expand(int n, long long nnz, const int* rp, const int* ci, const double* v,
                          const unsigned char* strong, const signed char* cf, const double* sum_strong,
                          const int* cnt_c, const int* f2c, const long long* cand_ptr, unsigned long long* keys,
                          double* vals, hipStream_t s) {
  if (wave_rows(n, nnz))
    launch_waves(k_interp_expand_wave, n, s, n, rp, ci, v, strong, cf, sum_strong, cnt_c, f2c, cand_ptr, keys, vals);
  else
    launch_rows(k_interp_expand, n, s, n, rp, ci, v, strong, cf, sum_strong, cnt_c, f2c, cand_ptr, keys, vals);
}
void launch_interp_reduce(int n, const int* rp, const int* ci, const signed char* cf, const int* f2c,
                          const long long* cand_ptr, const unsigned long long* keys, const double* vals,
                          double trunc, int pmax, int* out_cols, double* out_vals, int* pcnt, hipStream_t s) {
  launch_rows(k_interp_reduce, n, s, n, rp, ci, cf, f2c, cand_ptr, keys, vals, trunc, pmax, out_cols, out_vals, pcnt);
}
void launch_interp_gather(int n, const long long* cand_ptr, const int* cols_tmp, const double* vals_tmp,
                          const int* prp, int* pci, double* pv, hipStream_t s) {
  launch_rows(k_interp_gather, n, s, n, cand_ptr, cols_tmp, vals_tmp, prp, pci, pv);
}
void launch_col_count(int nnz, const int* ci, int* cnt, hipStream_t s) { launch_rows(k_col_count, nnz, s, nnz, ci, cnt); }
void launch_transpose_fill(int n_rows, const int* rp, const int* ci, const int* rrp, int* cursor,
                           unsigned long long* keys, hipStream_t s) {
  launch_rows(k_transpose_fill, n_rows, s, n_rows, rp, ci, rrp, cursor, keys);
}
void launch_transpose_finish(int nnz, const unsigned long long* keys, const double* src_v, int* rci, double* rv,
                             hipStream_t s) {
  launch_rows(k_transpose_finish, nnz, s, nnz, keys, src_v, rci, rv);
}
void launch_spgemm_count(int n, const int* lrp, const int* lci, const int* mrp, int* cnt, hipStream_t s) {
  launch_rows(k_spgemm_count, n, s, n, lrp, lci, mrp, cnt);
}
void launch_spgemm_expand(int n, long long left_nnz, long long products, const int* lrp, const int* lci, const int* mrp, const int* mci,
                          const long long* prod_ptr, unsigned long long* keys, int* ex_e, int* ex_f, hipStream_t s) {
  if (wave_rows_expand(n, left_nnz, products))
    launch_waves(k_spgemm_expand, n, s, n, lrp, lci, mrp, mci, prod_ptr, keys, ex_e, ex_f);
  else
    launch_rows(k_spgemm_expand_rows, n, s, n, lrp, lci, mrp, mci, prod_ptr, keys, ex_e, ex_f);
}
void launch_spgemm_unique_count(int n, const long long* prod_ptr, const unsigned long long* keys, int* cnt,
                                hipStream_t s) {
  launch_waves(k_spgemm_unique, n, s, n, prod_ptr, keys, cnt);
}
void launch_spgemm_finalize(int n, const long long* prod_ptr, const unsigned long long* keys, const int* crp,
                            const int* ex_e, const int* ex_f, int* cci, long long* seg, int* pe, int* pf,
                            long long total_products, int nnz_c, hipStream_t s) {
  launch_waves(k_spgemm_finalize, n, s, n, prod_ptr, keys, crp, ex_e, ex_f, cci, seg, pe, pf, total_products, nnz_c);
}
int spgemm_group_for(long long terms, int nnz_c) {
  if (nnz_c <= 0) return 1;
  const double mean = (double)terms / (double)nnz_c;
  int g = 1;
  while (g < 8 && g * 4 < mean) g <<= 1;
  return g;
}

void launch_spgemm_numeric(int nnz_c, const long long* seg, const int* pe, const int* pf, const double* lv,
                           const double* mv, double* cv, int G, hipStream_t s) {
  if (nnz_c <= 0) return;
  const int grid = grid_for((long long)nnz_c * G, kBlock, 1 << 16);
  switch (G) {
    case 1: k_spgemm_numeric<1><<<grid, kBlock, 0, s>>>(nnz_c, seg, pe, pf, lv, mv, cv); break;
    case 2: k_spgemm_numeric<2><<<grid, kBlock, 0, s>>>(nnz_c, seg, pe, pf, lv, mv, cv); break;
    case 4: k_spgemm_numeric<4><<<grid, kBlock, 0, s>>>(nnz_c, seg, pe, pf, lv, mv, cv); break;
    default: k_spgemm_numeric<8><<<grid, kBlock, 0, s>>>(nnz_c, seg, pe, pf, lv, mv, cv); break;
  }
  IMPES_LAUNCH_CHECK();
}

}  // namespace impes
