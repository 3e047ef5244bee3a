// pool.cpp -- per-thread cache of device memory blocks (see host_common.h).
#include <iterator>
#include <map>

#include "host_common.h"

namespace pm {
namespace {

struct Pool {
  std::multimap<size_t, void *> free_blocks;
  struct Kept {
    size_t cls = 0;
    void *p = nullptr;
  } slots[kSlotCount];
  ~Pool() {
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    for (auto &k : slots)
      if (k.p) (void)hipFree(k.p);
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

// Last in, first out within a class: a call that allocates and frees in the
// same order as the previous one gets the same blocks back (which also lets
// the optimiser graph of that call be replayed).
void *dev_acquire(size_t cls) {
  auto &fb = pool().free_blocks;
  const auto range = fb.equal_range(cls);
  if (range.first != range.second) {
    const auto it = std::prev(range.second);
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

void *dev_acquire_slot(int slot, size_t cls) {
  auto &k = pool().slots[slot];
  if (k.p && k.cls == cls) {
    void *p = k.p;
    k.p = nullptr;
    return p;
  }
  if (k.p) dev_recycle(k.p, k.cls);
  k.p = nullptr;
  return dev_acquire(cls);
}

void dev_release_slot(int slot, void *p, size_t cls) {
  auto &k = pool().slots[slot];
  if (k.p) {  // the slot already holds a block: this one joins the general cache
    dev_recycle(p, cls);
    return;
  }
  k.p = p;
  k.cls = cls;
}

void dev_trim() {
  auto &pl = pool();
  for (auto &k : pl.slots)
    if (k.p) {
      pl.free_blocks.emplace(k.cls, k.p);
      k.p = nullptr;
    }
  auto &fb = pl.free_blocks;
  if (fb.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto &kv : fb) (void)hipFree(kv.second);
  fb.clear();
}

}  // namespace pm

extern "C" void pacmap_release_cached_memory(void) {
  pm::graph_cache_reset();
  pm::dev_trim();
}
