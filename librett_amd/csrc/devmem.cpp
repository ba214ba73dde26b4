// Pluggable device allocator (see devmem.h) and its C ABI.
#include "devmem.h"

#include <mutex>
#include <unordered_map>

#include "../../include/librett_c.h"
#include "selector.h"

namespace lrt {
namespace {

struct Owner {
  librett_c_free_fn free_fn;
  void* ctx;
};

struct Allocator {
  std::mutex mu;
  librett_c_alloc_fn alloc_fn = nullptr;
  librett_c_free_fn free_fn = nullptr;
  void* ctx = nullptr;
  std::unordered_map<void*, Owner> owners;  // pointers made by a caller's allocator
};

Allocator& alloc_state() {
  static Allocator* a = new Allocator;
  return *a;
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
  Allocator& A = alloc_state();
  librett_c_alloc_fn fa;
  librett_c_free_fn ff;
  void* ctx;
  {
    std::lock_guard<std::mutex> lock(A.mu);
    fa = A.alloc_fn;
    ff = A.free_fn;
    ctx = A.ctx;
  }
  if (!fa) return hipMalloc(p, bytes);
  void* q = fa(bytes == 0 ? 1 : bytes, ctx);
  if (!q) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  std::lock_guard<std::mutex> lock(A.mu);
  A.owners[q] = Owner{ff, ctx};
  *p = q;
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  Allocator& A = alloc_state();
  Owner o{nullptr, nullptr};
  bool mine = false;
  {
    std::lock_guard<std::mutex> lock(A.mu);
    auto it = A.owners.find(p);
    if (it != A.owners.end()) {
      o = it->second;
      mine = true;
      A.owners.erase(it);
    }
  }
  if (mine)
    o.free_fn(p, o.ctx);
  else
    (void)hipFree(p);
}

}  // namespace lrt

extern "C" {

int librett_c_set_allocator(librett_c_alloc_fn alloc, librett_c_free_fn free_fn, void* ctx) {
  if ((alloc == nullptr) != (free_fn == nullptr)) return LIBRETT_C_INVALID_PARAMETER;
  lrt::Allocator& A = lrt::alloc_state();
  // The selector's inference buffers go back to the allocator that made
  // them before another one takes over.
  lrt::selector_release_buffers([&] {
    std::lock_guard<std::mutex> lock(A.mu);
    A.alloc_fn = alloc;
    A.free_fn = free_fn;
    A.ctx = ctx;
  });
  return LIBRETT_C_SUCCESS;
}

unsigned long long librett_c_allocator_live(void) {
  lrt::Allocator& A = lrt::alloc_state();
  std::lock_guard<std::mutex> lock(A.mu);
  return A.owners.size();
}

}  // extern "C"
