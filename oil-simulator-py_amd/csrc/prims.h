// Device primitives used by the AMG setup: exclusive scans, a segmented sort of
// 64-bit keys, and integer reductions. All deterministic (no order-dependent
// atomics in anything whose result is observable).
#pragma once

#include "common.h"

namespace impes {

// out[0..n] = exclusive scan of in[0..n-1], out[n] = total. Returns nothing;
// the total can be read from out[n]. Work buffers are allocated internally.
void scan_i32_to_i64(const int* in, long long* out, long long n, hipStream_t s);
void scan_i32_to_i32(const int* in, int* out, long long n, hipStream_t s);

// Sort every segment [seg[i], seg[i+1]) of `keys` ascending. Synchronises the
// stream once (bin sizes).
void segmented_sort_u64(unsigned long long* keys, const long long* seg, long long nseg, hipStream_t s);

// sum of an int32 array (synchronises, returns host value)
long long sum_i32(const int* x, long long n, hipStream_t s);
// count of int8 entries equal to v (synchronises)
long long count_i8(const signed char* x, long long n, signed char v, hipStream_t s);
// copy one value device->host (synchronises)
template <class T>
T read_scalar(const T* dev, hipStream_t s) {
  T h{};
  IMPES_HIP_CHECK(hipMemcpyAsync(&h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
  return h;
}

}  // namespace impes
