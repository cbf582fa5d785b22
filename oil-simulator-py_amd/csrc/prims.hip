// Scans, segmented sort and small reductions for the AMG setup phase.
#include "prims.h"

namespace impes {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kTile = kScanBlock * kScanItems;

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

// exclusive block scan; returns the exclusive prefix of v and writes the block total
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T& total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  constexpr int nw = kScanBlock / kWave;
  T inc = wave_incl_scan(v);
  __syncthreads();
  if (lane == kWave - 1) lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < nw; ++w) { T t = lds[w]; lds[w] = run; run += t; }
    lds[nw] = run;
  }
  __syncthreads();
  T res = lds[wid] + inc - v;
  total = lds[nw];
  __syncthreads();
  return res;
}

template <class TO>
__global__ __launch_bounds__(kScanBlock) void k_tile_sums(const int* __restrict__ in, long long n,
                                                          TO* __restrict__ bsum) {
  __shared__ TO lds[kScanBlock / kWave + 1];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kScanItems;
  TO s = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    long long i = base + t;
    if (i < n) s += (TO)in[i];
  }
  TO tot;
  (void)block_excl_scan<TO>(s, lds, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb tile sums in place, total -> *total
template <class TO>
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(TO* __restrict__ bsum, long long nb, TO* __restrict__ total) {
  __shared__ TO lds[kScanBlock / kWave + 1];
  TO carry = 0;
  for (long long c0 = 0; c0 < nb; c0 += kScanBlock) {
    long long i = c0 + threadIdx.x;
    TO v = (i < nb) ? bsum[i] : (TO)0;
    TO tot;
    TO ex = block_excl_scan<TO>(v, lds, tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

template <class TO>
__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const int* __restrict__ in, long long n,
                                                          const TO* __restrict__ bsum, TO* __restrict__ out) {
  __shared__ TO lds[kScanBlock / kWave + 1];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kScanItems;
  TO vals[kScanItems];
  TO s = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    long long i = base + t;
    vals[t] = (i < n) ? (TO)in[i] : (TO)0;
    s += vals[t];
  }
  TO tot;
  TO ex = block_excl_scan<TO>(s, lds, tot) + bsum[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    long long i = base + t;
    if (i < n) out[i] = ex;
    ex += vals[t];
  }
}

template <class TO>
void scan_impl(const int* in, TO* out, long long n, hipStream_t s) {
  if (n <= 0) {
    IMPES_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(TO), s));
    return;
  }
  const long long nb = (n + kTile - 1) / kTile;
  DevBuf<TO> bsum;
  bsum.alloc((size_t)nb);
  k_tile_sums<TO><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, bsum.p);
  IMPES_LAUNCH_CHECK();
  k_scan_tiles<TO><<<1, kScanBlock, 0, s>>>(bsum.p, nb, out + n);
  IMPES_LAUNCH_CHECK();
  k_tile_scan<TO><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, bsum.p, out);
  IMPES_LAUNCH_CHECK();
  // bsum must outlive the kernels: synchronise before it is freed.
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- segmented sort
// One wavefront per segment, keys staged in LDS and bitonic-sorted over the
// next power of two >= segment length (padding with all-ones keys).
template <int CAP>
__global__ __launch_bounds__(kWave) void k_segsort_lds(unsigned long long* __restrict__ keys,
                                                        const long long* __restrict__ seg, long long nseg) {
  __shared__ unsigned long long sk[CAP];
  for (long long sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const long long b = seg[sg];
    const int len = (int)(seg[sg + 1] - b);
    if (len <= 1 || len > CAP) continue;
    int p2 = 2;
    while (p2 < len) p2 <<= 1;
    for (int t = threadIdx.x; t < p2; t += kWave) sk[t] = (t < len) ? keys[b + t] : ~0ull;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = threadIdx.x; t < p2; t += kWave) {
          const int partner = t ^ stride;
          if (partner > t) {
            const bool asc = (t & size) == 0;
            unsigned long long a = sk[t], c = sk[partner];
            if ((a > c) == asc) { sk[t] = c; sk[partner] = a; }
          }
        }
        __syncthreads();
      }
    }
    for (int t = threadIdx.x; t < len; t += kWave) keys[b + t] = sk[t];
    __syncthreads();
  }
}

// Segments longer than the LDS capacity: one block walks them one at a time and
// bitonic-sorts through a global scratch buffer of the padded length.
__global__ __launch_bounds__(1024) void k_segsort_global(unsigned long long* __restrict__ keys,
                                                         const long long* __restrict__ seg, long long nseg,
                                                         int min_len, unsigned long long* __restrict__ scratch) {
  for (long long sg = 0; sg < nseg; ++sg) {
    const long long b = seg[sg];
    const long long len = seg[sg + 1] - b;
    if (len < min_len) continue;
    long long p2 = 2;
    while (p2 < len) p2 <<= 1;
    for (long long t = threadIdx.x; t < p2; t += blockDim.x) scratch[t] = (t < len) ? keys[b + t] : ~0ull;
    __syncthreads();
    for (long long size = 2; size <= p2; size <<= 1) {
      for (long long stride = size >> 1; stride > 0; stride >>= 1) {
        for (long long t = threadIdx.x; t < p2; t += blockDim.x) {
          const long long partner = t ^ stride;
          if (partner > t) {
            const bool asc = (t & size) == 0;
            unsigned long long a = scratch[t], c = scratch[partner];
            if ((a > c) == asc) { scratch[t] = c; scratch[partner] = a; }
          }
        }
        __syncthreads();
      }
    }
    for (long long t = threadIdx.x; t < len; t += blockDim.x) keys[b + t] = scratch[t];
    __syncthreads();
  }
}

__global__ void k_seg_maxlen(const long long* __restrict__ seg, long long nseg, long long* __restrict__ part) {
  __shared__ long long lds[256 / kWave];
  long long m = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nseg; i += (long long)gridDim.x * blockDim.x) {
    long long l = seg[i + 1] - seg[i];
    m = l > m ? l : m;
  }
  m = block_max<long long>(m, lds, 0);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_sum_i32(const int* __restrict__ x, long long n, long long* __restrict__ part) {
  __shared__ long long lds[256 / kWave];
  long long m = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m += x[i];
  m = block_sum<long long>(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_count_i8(const signed char* __restrict__ x, long long n, signed char v, long long* __restrict__ part) {
  __shared__ long long lds[256 / kWave];
  long long m = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m += (x[i] == v) ? 1 : 0;
  m = block_sum<long long>(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

constexpr int kPartBlocks = 256;

long long fold_host(const DevBuf<long long>& part, bool is_max, hipStream_t s) {
  long long h[kPartBlocks];
  IMPES_HIP_CHECK(hipMemcpyAsync(h, part.p, sizeof(h), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  long long r = 0;
  for (int i = 0; i < kPartBlocks; ++i) r = is_max ? (h[i] > r ? h[i] : r) : r + h[i];
  return r;
}

}  // namespace

void scan_i32_to_i64(const int* in, long long* out, long long n, hipStream_t s) { scan_impl<long long>(in, out, n, s); }
void scan_i32_to_i32(const int* in, int* out, long long n, hipStream_t s) { scan_impl<int>(in, out, n, s); }

void segmented_sort_u64(unsigned long long* keys, const long long* seg, long long nseg, long long max_len,
                        hipStream_t s) {
  if (nseg <= 0 || max_len <= 1) return;
  const int grid = (int)(nseg < 65536 ? nseg : 65536);
  if (max_len <= 64) {
    k_segsort_lds<64><<<grid, kWave, 0, s>>>(keys, seg, nseg);
  } else if (max_len <= 256) {
    k_segsort_lds<256><<<grid, kWave, 0, s>>>(keys, seg, nseg);
  } else if (max_len <= 1024) {
    k_segsort_lds<1024><<<grid, kWave, 0, s>>>(keys, seg, nseg);
  } else {
    k_segsort_lds<4096><<<grid, kWave, 0, s>>>(keys, seg, nseg);
  }
  IMPES_LAUNCH_CHECK();
  if (max_len > 4096) {
    long long p2 = 2;
    while (p2 < max_len) p2 <<= 1;
    DevBuf<unsigned long long> scratch;
    scratch.alloc((size_t)p2);
    k_segsort_global<<<1, 1024, 0, s>>>(keys, seg, nseg, 4097, scratch.p);
    IMPES_LAUNCH_CHECK();
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
  }
}

long long max_segment_length(const long long* seg, long long nseg, hipStream_t s) {
  if (nseg <= 0) return 0;
  DevBuf<long long> part;
  part.alloc(kPartBlocks);
  k_seg_maxlen<<<kPartBlocks, 256, 0, s>>>(seg, nseg, part.p);
  IMPES_LAUNCH_CHECK();
  return fold_host(part, true, s);
}

long long sum_i32(const int* x, long long n, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf<long long> part;
  part.alloc(kPartBlocks);
  k_sum_i32<<<kPartBlocks, 256, 0, s>>>(x, n, part.p);
  IMPES_LAUNCH_CHECK();
  return fold_host(part, false, s);
}

long long count_i8(const signed char* x, long long n, signed char v, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf<long long> part;
  part.alloc(kPartBlocks);
  k_count_i8<<<kPartBlocks, 256, 0, s>>>(x, n, v, part.p);
  IMPES_LAUNCH_CHECK();
  return fold_host(part, false, s);
}

}  // namespace impes
