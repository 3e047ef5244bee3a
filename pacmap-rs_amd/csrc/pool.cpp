// pool.cpp -- per-thread cache of device memory blocks and the staged
// host-to-device upload (see host_common.h).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <vector>
#include <condition_variable>
#include <iterator>
#include <map>
#include <mutex>
#include <thread>

#include <functional>
#include <stdexcept>

#include "host_common.h"

namespace pm {

// Set by an exit hook registered on first use.  Thread-exit destructors run
// before the process's static destructors (the HIP runtime's among them), so
// they may release HIP resources; once the hook has run, static destruction
// is under way and they leave them to the end of the process.
namespace {
std::atomic<bool> g_exiting{false};
std::once_flag g_exit_hook;
}  // namespace
void note_runtime_use() {
  std::call_once(g_exit_hook, [] { std::atexit([] { g_exiting.store(true); }); });
}
bool process_exiting() { return g_exiting.load(); }

namespace {

// One thread's cache of idle device blocks.  Every pool is listed in a
// process-wide registry, so that an out-of-memory retry or
// pacmap_release_cached_memory frees the idle blocks of every thread; each
// pool has its own lock (uncontended on the owner's calls).
struct Pool;
struct Registry {
  std::mutex mu;
  std::vector<Pool *> pools;
};
Registry &registry() {
  static Registry *r = new Registry;  // never destroyed: pools unregister at thread exit
  return *r;
}

struct Pool {
  std::mutex mu;
  std::multimap<size_t, void *> free_blocks;
  struct Kept {
    size_t cls = 0;
    void *p = nullptr;
  } slots[kSlotCount];
  Pool() {
    note_runtime_use();
    Registry &r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    r.pools.push_back(this);
  }
  // moves every idle block (slots included) out of the pool
  void take_all(std::vector<void *> &out) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &kv : free_blocks) out.push_back(kv.second);
    free_blocks.clear();
    for (auto &k : slots)
      if (k.p) {
        out.push_back(k.p);
        k.p = nullptr;
      }
  }
  ~Pool() {
    {
      Registry &r = registry();
      std::lock_guard<std::mutex> lk(r.mu);
      r.pools.erase(std::remove(r.pools.begin(), r.pools.end(), this), r.pools.end());
    }
    if (process_exiting()) return;
    std::vector<void *> blocks;
    take_all(blocks);
    for (void *p : blocks) (void)hipFree(p);
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
  Pool &pl = pool();
  {
    std::lock_guard<std::mutex> lk(pl.mu);
    auto &fb = pl.free_blocks;
    const auto range = fb.equal_range(cls);
    if (range.first != range.second) {
      const auto it = std::prev(range.second);
      void *p = it->second;
      fb.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {  // out of memory: give every thread's idle blocks back and retry once
    (void)hipGetLastError();
    dev_trim();
    e = hipMalloc(&p, cls);
  }
  if (e != hipSuccess) fail_gpu(e, "hipMalloc", __FILE__, __LINE__);
  return p;
}

void dev_recycle(void *p, size_t cls) {
  Pool &pl = pool();
  std::lock_guard<std::mutex> lk(pl.mu);
  pl.free_blocks.emplace(cls, p);
}

void *dev_acquire_slot(int slot, size_t cls) {
  Pool &pl = pool();
  void *old = nullptr;
  size_t old_cls = 0;
  {
    std::lock_guard<std::mutex> lk(pl.mu);
    auto &k = pl.slots[slot];
    if (k.p && k.cls == cls) {
      void *p = k.p;
      k.p = nullptr;
      return p;
    }
    old = k.p;
    old_cls = k.cls;
    k.p = nullptr;
  }
  if (old) dev_recycle(old, old_cls);
  return dev_acquire(cls);
}

void dev_release_slot(int slot, void *p, size_t cls) {
  Pool &pl = pool();
  std::lock_guard<std::mutex> lk(pl.mu);
  auto &k = pl.slots[slot];
  if (k.p) {  // the slot already holds a block: this one joins the general cache
    pl.free_blocks.emplace(cls, p);
    return;
  }
  k.p = p;
  k.cls = cls;
}

// Frees the idle blocks of every thread's pool (blocks in use belong to
// DevBufs, not to the pools).  The device is synchronised first: a block
// can be released while work queued on its owner's stream still uses it.
void dev_trim() {
  std::vector<void *> blocks;
  {
    Registry &r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    for (Pool *pl : r.pools) pl->take_all(blocks);
  }
  if (blocks.empty()) return;
  (void)hipDeviceSynchronize();
  for (void *p : blocks) (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// Staged uploads.  A copy from pageable host memory runs at about half the
// DMA rate of pinned memory (the runtime stages it through one thread), so
// large uploads go through two pinned buffers: host threads copy the next
// chunk into one buffer while the DMA engine moves the previous one from the
// other.  Each buffer is rewritten only after the event recorded behind its
// last DMA has completed.
// ---------------------------------------------------------------------------
namespace {

// Persistent host threads that split a memcpy between them (created on first
// use; one copy at a time, callers are serialised).
class CopyThreads {
 public:
  static CopyThreads &get() {
    static CopyThreads c;
    return c;
  }
  void copy(void *dst, const void *src, size_t len) {
    std::lock_guard<std::mutex> serial(call_);
    std::unique_lock<std::mutex> lk(m_);
    dst_ = static_cast<char *>(dst);
    src_ = static_cast<const char *>(src);
    len_ = len;
    pending_ = static_cast<int>(th_.size());
    ++gen_;
    cv_.notify_all();
    lk.unlock();
    part(0, static_cast<int>(th_.size()) + 1);  // the caller copies the first share
    lk.lock();
    done_.wait(lk, [&] { return pending_ == 0; });
  }
  ~CopyThreads() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }

 private:
  CopyThreads() {
    const int n = std::max(1, std::min(8, host_threads()) - 1);
    for (int i = 0; i < n; ++i) th_.emplace_back([this, i] { run(i + 1); });
  }
  void part(int id, int parts) {
    const size_t b = len_ * id / parts, e = len_ * (id + 1) / parts;
    if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
  }
  void run(int id) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      const int parts = static_cast<int>(th_.size()) + 1;
      lk.unlock();
      part(id, parts);
      lk.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex call_, m_;
  std::condition_variable cv_, done_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t len_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

constexpr size_t kStageChunk = size_t{32} << 20;  // bytes per pinned buffer
constexpr size_t kStageMin = size_t{8} << 20;     // smaller uploads go directly

struct Staging {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t order = nullptr;  // orders a copy stream after the calling stream
  ~Staging() {
    if (!process_exiting()) release();
  }
  void release() {
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) {
        (void)hipEventSynchronize(ev[b]);
        (void)hipEventDestroy(ev[b]);
      }
      if (buf[b]) (void)hipHostFree(buf[b]);
      buf[b] = nullptr;
      ev[b] = nullptr;
    }
    if (order) (void)hipEventDestroy(order);
    order = nullptr;
  }
};
Staging &staging() {
  thread_local Staging st;  // kept for the thread's calls, freed when the thread exits
  note_runtime_use();
  return st;
}

}  // namespace

void upload_bytes(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (bytes < kStageMin) {
    PM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  Staging &st = staging();
  for (int b = 0; b < 2; ++b) {
    if (!st.buf[b]) PM_HIP(hipHostMalloc(&st.buf[b], kStageChunk, 0));
    if (!st.ev[b]) PM_HIP(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
  }
  const char *sp = static_cast<const char *>(src);
  char *dp = static_cast<char *>(dst);
  size_t k = 0;
  for (size_t off = 0; off < bytes; off += kStageChunk, ++k) {
    const int b = static_cast<int>(k & 1);
    const size_t len = std::min(kStageChunk, bytes - off);
    PM_HIP(hipEventSynchronize(st.ev[b]));  // the last DMA out of this buffer is done
    CopyThreads::get().copy(st.buf[b], sp + off, len);
    PM_HIP(hipMemcpyAsync(dp + off, st.buf[b], len, hipMemcpyHostToDevice, s));
    PM_HIP(hipEventRecord(st.ev[b], s));
  }
}

void upload_rows_pieces(void *dst, const void *src, uint64_t rows, size_t row_bytes, uint64_t piece_rows,
                        hipStream_t s, hipStream_t cs, const std::function<void(uint32_t, hipEvent_t)> &on_piece) {
  if (rows == 0) return;
  if (piece_rows == 0 || piece_rows * row_bytes > kStageChunk) throw std::invalid_argument("upload piece too large");
  Staging &st = staging();
  for (int b = 0; b < 2; ++b) {
    if (!st.buf[b]) PM_HIP(hipHostMalloc(&st.buf[b], kStageChunk, 0));
    if (!st.ev[b]) PM_HIP(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
  }
  // the copy stream starts after the work already queued on s (dst may be a
  // block that work still reads)
  if (!st.order) PM_HIP(hipEventCreateWithFlags(&st.order, hipEventDisableTiming));
  PM_HIP(hipEventRecord(st.order, s));
  PM_HIP(hipStreamWaitEvent(cs, st.order, 0));
  const char *sp = static_cast<const char *>(src);
  char *dp = static_cast<char *>(dst);
  const uint64_t pieces = (rows + piece_rows - 1) / piece_rows;
  for (uint64_t k = 0; k < pieces; ++k) {
    const int b = static_cast<int>(k & 1);
    const size_t off = static_cast<size_t>(k * piece_rows) * row_bytes;
    const size_t len = static_cast<size_t>(std::min(piece_rows, rows - k * piece_rows)) * row_bytes;
    PM_HIP(hipEventSynchronize(st.ev[b]));  // the last DMA out of this buffer is done
    CopyThreads::get().copy(st.buf[b], sp + off, len);
    PM_HIP(hipMemcpyAsync(dp + off, st.buf[b], len, hipMemcpyHostToDevice, cs));
    PM_HIP(hipEventRecord(st.ev[b], cs));
    on_piece(static_cast<uint32_t>(k), st.ev[b]);
  }
}

void staging_release() { staging().release(); }

}  // namespace pm

extern "C" void pacmap_release_cached_memory(void) {
  pm::graph_cache_reset();
  pm::dev_trim();
  pm::staging_release();
  pm::mlp_transform_cache_release();
}
