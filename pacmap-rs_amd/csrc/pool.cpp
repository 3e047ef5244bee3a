// pool.cpp -- per-thread cache of device memory blocks (see host_common.h).
#include <map>

#include "host_common.h"

namespace pm {
namespace {

struct Pool {
  std::multimap<size_t, void *> free_blocks;
  ~Pool() {
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
  }
};

Pool &pool() {
  thread_local Pool p;
  return p;
}

}  // namespace

// Classes: multiples of 4 KiB up to 1 MiB, then eight steps per power of
// two (at most 12.5 % above the request).
size_t dev_size_class(size_t bytes) {
  if (bytes <= (1u << 20)) return (bytes + 4095) / 4096 * 4096;
  size_t p = 1u << 20;
  while (p * 2 < bytes) p *= 2;
  const size_t step = p / 8;
  return (bytes + step - 1) / step * step;
}

void *dev_acquire(size_t cls) {
  auto &fb = pool().free_blocks;
  auto it = fb.find(cls);
  if (it != fb.end()) {
    void *p = it->second;
    fb.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {  // out of memory: give the cache back and retry once
    (void)hipGetLastError();
    dev_trim();
    e = hipMalloc(&p, cls);
  }
  if (e != hipSuccess) fail_gpu(e, "hipMalloc", __FILE__, __LINE__);
  return p;
}

void dev_recycle(void *p, size_t cls) { pool().free_blocks.emplace(cls, p); }

void dev_trim() {
  auto &fb = pool().free_blocks;
  if (fb.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto &kv : fb) (void)hipFree(kv.second);
  fb.clear();
}

}  // namespace pm

extern "C" void pacmap_release_cached_memory(void) { pm::dev_trim(); }
