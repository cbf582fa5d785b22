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
  sync_before_release(s);  // bsum must outlive the kernels
}

// ---------------------------------------------------------------- segmented sort
// Segments are binned by length (k_seg_bin), and each bin is sorted by a
// kernel sized for it: a thread per segment of <= 16 keys (registers), else a
// workgroup per segment with the keys staged in LDS and bitonic-sorted over the
// next power of two >= the segment length (padding with all-ones keys), from
// one wavefront for <= 64 keys up to 1024 threads for <= 8192 keys (64 KiB of
// LDS); longer segments sort the same way through a global scratch region per
// workgroup. Bins only choose the kernel: each
// segment is sorted on its own, so the result does not depend on the binning.
constexpr int kSortBins = 7;  // <= 16, 64, 256, 1024, 4096, 8192 keys, and longer

__device__ __forceinline__ int sort_bin(long long len) {
  return len <= 16 ? 0 : len <= 64 ? 1 : len <= 256 ? 2 : len <= 1024 ? 3 : len <= 4096 ? 4 : len <= 8192 ? 5 : 6;
}

// Segments of at most 16 keys: a thread per segment, bitonic network over 16
// registers (padding with all-ones keys).
__global__ void k_segsort_regs(unsigned long long* __restrict__ keys, const long long* __restrict__ seg,
                               const int* __restrict__ list, int count) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < count;
       q += (long long)gridDim.x * blockDim.x) {
    const long long sg = list[q];
    const long long b = seg[sg];
    const int len = (int)(seg[sg + 1] - b);
    unsigned long long r[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) r[t] = (t < len) ? keys[b + t] : ~0ull;
#pragma unroll
    for (int size = 2; size <= 16; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const int partner = t ^ stride;
          if (partner > t) {
            const bool asc = (t & size) == 0;
            const unsigned long long a = r[t], c = r[partner];
            if ((a > c) == asc) { r[t] = c; r[partner] = a; }
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (t < len) keys[b + t] = r[t];
  }
}

__global__ void k_fill_zero(int* __restrict__ counts, int n, long long* __restrict__ m) {
  if (threadIdx.x < n) counts[threadIdx.x] = 0;
  if (threadIdx.x == 0) *m = 0;
}

// Bin lists: each wavefront claims its slots with one atomic per bin present
// in it (a single counter per bin would otherwise take one atomic per segment).
__global__ void k_seg_bin(const long long* __restrict__ seg, long long nseg, int* __restrict__ counts,
                          long long* __restrict__ long_max, int* __restrict__ lists) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  for (long long i0 = blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~(kWave - 1)); i0 < nseg;
       i0 += (long long)gridDim.x * blockDim.x) {
    const long long i = i0 + lane;
    int bin = -1;
    long long len = 0;
    if (i < nseg) {
      len = seg[i + 1] - seg[i];
      if (len > 1) bin = sort_bin(len);
    }
    for (int b = 0; b < kSortBins; ++b) {
      const unsigned long long m = __ballot(bin == b);
      if (!m) continue;
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&counts[b], __popcll(m));
      base = __shfl(base, leader, kWave);
      if (bin == b) lists[(long long)b * nseg + base + __popcll(m & below)] = (int)i;
    }
    if (bin == kSortBins - 1) atomicMax((unsigned long long*)long_max, (unsigned long long)len);
  }
}

template <class K>
__device__ __forceinline__ void bitonic_block(K* sk, long long p2) {
  for (long long size = 2; size <= p2; size <<= 1) {
    for (long long stride = size >> 1; stride > 0; stride >>= 1) {
      for (long long t = threadIdx.x; t < p2; t += blockDim.x) {
        const long long partner = t ^ stride;
        if (partner > t) {
          const bool asc = (t & size) == 0;
          const K a = sk[t], c = sk[partner];
          if ((a > c) == asc) { sk[t] = c; sk[partner] = a; }
        }
      }
      __syncthreads();
    }
  }
}

template <int CAP, int NT>
__global__ __launch_bounds__(NT) void k_segsort_bin(unsigned long long* __restrict__ keys,
                                                    const long long* __restrict__ seg, const int* __restrict__ list,
                                                    int count) {
  __shared__ unsigned long long sk[CAP];
  for (int q = blockIdx.x; q < count; q += gridDim.x) {
    const long long sg = list[q];
    const long long b = seg[sg];
    const int len = (int)(seg[sg + 1] - b);
    int p2 = 2;
    while (p2 < len) p2 <<= 1;
    for (int t = threadIdx.x; t < p2; t += NT) sk[t] = (t < len) ? keys[b + t] : ~0ull;
    __syncthreads();
    bitonic_block(sk, (long long)p2);
    for (int t = threadIdx.x; t < len; t += NT) keys[b + t] = sk[t];
    __syncthreads();
  }
}

// Segments longer than the LDS bins: workgroup g sorts list entries g, g + grid,
// ... in scratch[g * p2max, (g + 1) * p2max).
__global__ __launch_bounds__(1024) void k_segsort_global(unsigned long long* __restrict__ keys,
                                                         const long long* __restrict__ seg,
                                                         const int* __restrict__ list, int count, long long p2max,
                                                         unsigned long long* __restrict__ scratch) {
  unsigned long long* sk = scratch + blockIdx.x * p2max;
  for (int q = blockIdx.x; q < count; q += gridDim.x) {
    const long long sg = list[q];
    const long long b = seg[sg];
    const long long len = seg[sg + 1] - b;
    long long p2 = 2;
    while (p2 < len) p2 <<= 1;
    for (long long t = threadIdx.x; t < p2; t += blockDim.x) sk[t] = (t < len) ? keys[b + t] : ~0ull;
    __syncthreads();
    bitonic_block(sk, p2);
    for (long long t = threadIdx.x; t < len; t += blockDim.x) keys[b + t] = sk[t];
    __syncthreads();
  }
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

void segmented_sort_u64(unsigned long long* keys, const long long* seg, long long nseg, hipStream_t s) {
  if (nseg <= 0) return;
  IMPES_REQUIRE(nseg < (1LL << 31), "segmented sort: too many segments");
  DevBuf<int> counts, lists;
  DevBuf<long long> long_max;
  counts.alloc(kSortBins);
  long_max.alloc(1);
  lists.alloc((size_t)kSortBins * nseg);
  k_fill_zero<<<1, 64, 0, s>>>(counts.p, kSortBins, long_max.p);
  IMPES_LAUNCH_CHECK();
  k_seg_bin<<<grid_for(nseg, 256, 2048), 256, 0, s>>>(seg, nseg, counts.p, long_max.p, lists.p);
  IMPES_LAUNCH_CHECK();
  struct { int c[kSortBins]; long long m; } h;
  IMPES_HIP_CHECK(hipMemcpyAsync(h.c, counts.p, sizeof(h.c), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipMemcpyAsync(&h.m, long_max.p, sizeof(h.m), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  auto grid = [](int c) { return c < 65536 ? c : 65536; };
  const int* L = lists.p;
  if (h.c[0]) k_segsort_regs<<<grid_for(h.c[0], 256, 8192), 256, 0, s>>>(keys, seg, L, h.c[0]);
  if (h.c[1]) k_segsort_bin<64, 64><<<grid(h.c[1]), 64, 0, s>>>(keys, seg, L + nseg, h.c[1]);
  if (h.c[2]) k_segsort_bin<256, 128><<<grid(h.c[2]), 128, 0, s>>>(keys, seg, L + 2 * nseg, h.c[2]);
  if (h.c[3]) k_segsort_bin<1024, 256><<<grid(h.c[3]), 256, 0, s>>>(keys, seg, L + 3 * nseg, h.c[3]);
  if (h.c[4]) k_segsort_bin<4096, 512><<<grid(h.c[4]), 512, 0, s>>>(keys, seg, L + 4 * nseg, h.c[4]);
  if (h.c[5]) k_segsort_bin<8192, 1024><<<grid(h.c[5]), 1024, 0, s>>>(keys, seg, L + 5 * nseg, h.c[5]);
  IMPES_LAUNCH_CHECK();
  if (h.c[6]) {
    long long p2 = 2;
    while (p2 < h.m) p2 <<= 1;
    // at most 2^27 scratch keys (1 GiB) at a time
    long long g = (1LL << 27) / p2;
    g = g < 1 ? 1 : (g > h.c[6] ? h.c[6] : g);
    g = g > 1024 ? 1024 : g;
    DevBuf<unsigned long long> scratch;
    scratch.alloc((size_t)(g * p2));
    k_segsort_global<<<(unsigned)g, 1024, 0, s>>>(keys, seg, L + 6 * nseg, h.c[6], p2, scratch.p);
    IMPES_LAUNCH_CHECK();
  }
  sync_before_release(s);  // the bin lists and the scratch must outlive the kernels
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
