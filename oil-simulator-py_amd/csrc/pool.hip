// Device memory of hierarchy builds: slab sub-allocation and deferred release
// (SetupPool, PoolScope; declared in common.h). Host code only.
#include <map>
#include <mutex>

#include "common.h"

namespace impes {

struct SetupPool::Impl {
  struct Slab {
    size_t size = 0, used = 0;
    long live = 0;
  };
  std::mutex m;
  std::map<char*, Slab> slabs;  // by base address
  char* cur = nullptr;          // slab that new small blocks are carved from
  std::map<void*, size_t> large;        // live large blocks and their sizes
  std::multimap<size_t, void*> spare;   // large blocks released under the scope
  int open = 0;

  std::map<char*, Slab>::iterator slab_of(void* p) {
    auto it = slabs.upper_bound(static_cast<char*>(p));
    if (it == slabs.begin()) return slabs.end();
    --it;
    if (static_cast<char*>(p) >= it->first + it->second.size) return slabs.end();
    return it;
  }
  void free_slab(std::map<char*, Slab>::iterator it) {
    if (it->first == cur) cur = nullptr;
    (void)hipFree(it->first);
    slabs.erase(it);
  }
};

SetupPool::SetupPool() = default;

SetupPool::Impl* SetupPool::impl() {
  if (!impl_) impl_.reset(new Impl());
  return impl_.get();
}

SetupPool::~SetupPool() {
  if (!impl_) return;
  for (auto& b : impl_->spare) (void)hipFree(b.second);
  for (auto& b : impl_->large) (void)hipFree(b.first);
  for (auto& s : impl_->slabs) (void)hipFree(s.first);
}

void* SetupPool::alloc(size_t bytes) {
  Impl* I = impl();
  std::lock_guard<std::mutex> g(I->m);
  IMPES_REQUIRE(I->open > 0, "setup pool: allocation outside a scope");
  if (bytes > kPoolSmall) {
    auto it = I->spare.lower_bound(bytes);
    if (it != I->spare.end() && it->first <= 2 * bytes) {
      void* p = it->second;
      I->large[p] = it->first;
      I->spare.erase(it);
      return p;
    }
    void* p = nullptr;
    IMPES_HIP_CHECK(hipMalloc(&p, bytes));
    I->large[p] = bytes;
    return p;
  }
  const size_t r = (bytes + 255) & ~size_t(255);
  if (!I->cur || I->slabs[I->cur].used + r > I->slabs[I->cur].size) {
    char* base = nullptr;
    IMPES_HIP_CHECK(hipMalloc((void**)&base, kSlab));
    I->slabs[base] = Impl::Slab{kSlab, 0, 0};
    I->cur = base;
  }
  Impl::Slab& s = I->slabs[I->cur];
  char* p = I->cur + s.used;
  s.used += r;
  ++s.live;
  return p;
}

void SetupPool::release(void* p) {
  Impl* I = impl();
  std::lock_guard<std::mutex> g(I->m);
  auto lit = I->large.find(p);
  if (lit != I->large.end()) {
    if (I->open > 0) I->spare.emplace(lit->second, p);
    else (void)hipFree(p);
    I->large.erase(lit);
    return;
  }
  auto it = I->slab_of(p);
  if (it == I->slabs.end()) return;  // not a block of this pool
  --it->second.live;
  if (I->open == 0 && it->second.live == 0) I->free_slab(it);
}

void SetupPool::open() {
  Impl* I = impl();
  std::lock_guard<std::mutex> g(I->m);
  ++I->open;
}

void SetupPool::close() {
  Impl* I = impl();
  std::lock_guard<std::mutex> g(I->m);
  if (--I->open > 0) return;
  for (auto& b : I->spare) (void)hipFree(b.second);
  I->spare.clear();
  for (auto it = I->slabs.begin(); it != I->slabs.end();) {
    auto next = std::next(it);
    if (it->second.live == 0) I->free_slab(it);
    it = next;
  }
  I->cur = nullptr;  // slab space is not reused across scopes either
}

size_t SetupPool::slab_bytes() const {
  if (!impl_) return 0;
  std::lock_guard<std::mutex> g(impl_->m);
  size_t b = 0;
  for (auto& s : impl_->slabs) b += s.second.size;
  return b;
}

namespace {
thread_local std::shared_ptr<SetupPool> t_pool;
}

std::shared_ptr<SetupPool> current_setup_pool() { return t_pool; }

PoolScope::PoolScope(std::shared_ptr<SetupPool> pool) : pool_(std::move(pool)), prev_(t_pool) {
  pool_->open();
  t_pool = pool_;
}

PoolScope::~PoolScope() {
  t_pool = prev_;
  pool_->close();
}

}  // namespace impes
