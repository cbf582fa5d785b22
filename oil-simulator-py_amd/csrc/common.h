// Shared definitions for the IMPES HIP library (gfx950 / CDNA4 only).
//
// Conventions
//  * every extern "C" entry point returns int (0 = ok, <0 = error) and records a
//    message retrievable through impes_last_error() (thread-local);
//  * all device work goes on the caller's hipStream_t;
//  * indices are int32 (the largest operator handled, nnz(A0) = 6.94e6 at
//    100x100x100, is far below 2^31), values are double or float.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace impes {

constexpr int kWave = 64;  // CDNA wavefront width
constexpr int kXcd = 8;    // accelerator complex dies per MI355X; each has its own L2

// ---------------------------------------------------------------- errors
struct Error : public std::runtime_error {
  int code;
  explicit Error(const std::string& m, int c = -1) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define IMPES_HIP_CHECK(expr)                                                     \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      throw ::impes::Error(std::string("HIP error '") + hipGetErrorString(_e) +  \
                           "' at " + __FILE__ + ":" + std::to_string(__LINE__) + \
                           " in " #expr, -2);                                    \
    }                                                                             \
  } while (0)

#define IMPES_REQUIRE(cond, msg)                                   \
  do {                                                             \
    if (!(cond)) throw ::impes::Error(std::string(msg), -3);       \
  } while (0)

#define IMPES_LAUNCH_CHECK() IMPES_HIP_CHECK(hipGetLastError())

// Wraps a C-ABI body: converts exceptions into an error code + message.
template <class F>
int guarded(const char* fn, F&& body) {
  try {
    body();
    return 0;
  } catch (const Error& e) {
    set_last_error(std::string(fn) + ": " + e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(std::string(fn) + ": " + e.what());
    return -1;
  } catch (...) {
    set_last_error(std::string(fn) + ": unknown exception");
    return -1;
  }
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Grid size for a grid-stride loop over n items with `block` threads each,
// capped so a launch is large enough to fill 256 CUs but no larger.
inline int grid_for(long long n, int block, int cap = 8192) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// ---------------------------------------------------------------- device helpers
// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD b % 8).
// This renumbering gives every XCD one contiguous range of logical blocks, so
// neighbouring rows (stencil planes, the gather footprint of a CSR row range)
// share one L2 instead of eight.
__device__ __forceinline__ long long xcd_block() {
  const unsigned b = blockIdx.x, g = gridDim.x;
  const unsigned q = g / kXcd, r = g % kXcd, x = b % kXcd;
  return (long long)x * q + (x < r ? x : r) + b / kXcd;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// Sum within aligned groups of G lanes (G a power of two <= 64).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum; every thread gets the result. `lds` needs blockDim/64 slots.
template <class T>
__device__ __forceinline__ T block_sum(T v, T* lds) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T t = (threadIdx.x < nw) ? lds[threadIdx.x] : T(0);
  if (wid == 0) t = wave_sum(t);
  if (threadIdx.x == 0) lds[0] = t;
  __syncthreads();
  T r = lds[0];
  __syncthreads();
  return r;
}

template <class T>
__device__ __forceinline__ T block_max(T v, T* lds, T init) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T t = (threadIdx.x < nw) ? lds[threadIdx.x] : init;
  if (wid == 0) t = wave_max(t);
  if (threadIdx.x == 0) lds[0] = t;
  __syncthreads();
  T r = lds[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- setup memory
// A hierarchy build makes several hundred device allocations, and every hipFree
// also waits for the device. While a PoolScope is open on the calling thread,
// DevBuf allocations of at most kPoolSmall bytes are carved from 32 MiB slabs
// and larger ones are separate blocks. A small block released under the scope
// only lowers its slab's live count (slab space is never handed out twice); a
// large one becomes a spare that a later allocation of the scope may take (at
// most twice the size asked for). A spare's previous users were enqueued
// earlier, so the scope's owner must put all device work of the scope on one
// stream, and close the scope after synchronising it: closing frees the spares
// and every slab without live blocks. Outside a scope, blocks of the pool are
// released at once (hipFree, or the slab when it empties).
class SetupPool {
 public:
  static constexpr size_t kPoolSmall = size_t(1) << 20;
  static constexpr size_t kSlab = size_t(32) << 20;
  SetupPool();
  SetupPool(const SetupPool&) = delete;
  SetupPool& operator=(const SetupPool&) = delete;
  ~SetupPool();
  void* alloc(size_t bytes);
  void release(void* p);
  void open();
  void close();
  size_t slab_bytes() const;  // device memory held in slabs

 private:
  struct Impl;
  Impl* impl();
  std::unique_ptr<Impl> impl_;
};

// The pool that DevBuf allocations use on this thread (null outside a scope).
std::shared_ptr<SetupPool> current_setup_pool();

// Waits for s before buffers that its queued work uses are released. Under a
// PoolScope releases are deferred to the scope's end, so there it waits for nothing.
inline void sync_before_release(hipStream_t s) {
  if (!current_setup_pool()) IMPES_HIP_CHECK(hipStreamSynchronize(s));
}

class PoolScope {
 public:
  explicit PoolScope(std::shared_ptr<SetupPool> pool);
  ~PoolScope();
  PoolScope(const PoolScope&) = delete;
  PoolScope& operator=(const PoolScope&) = delete;

 private:
  std::shared_ptr<SetupPool> pool_, prev_;
};

// ---------------------------------------------------------------- device buffers
// Minimal RAII owner of a device allocation (hipMalloc, or a setup pool's block
// when allocated under a PoolScope); used by handles that own hierarchy /
// workspace memory for their lifetime.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  std::shared_ptr<SetupPool> pool;  // set when the block came from a setup pool
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), pool(std::move(o.pool)) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; pool = std::move(o.pool);
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (pool) pool->release(p);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    pool.reset();
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    std::shared_ptr<SetupPool> sp = current_setup_pool();
    if (sp) {
      p = static_cast<T*>(sp->alloc(count * sizeof(T)));
      pool = std::move(sp);
    } else {
      IMPES_HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    }
    n = count;
  }
  // grow-only allocation (keeps the buffer when already large enough)
  void reserve(size_t count) {
    if (count <= n && p) return;
    alloc(count);
  }
  T* get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace impes
