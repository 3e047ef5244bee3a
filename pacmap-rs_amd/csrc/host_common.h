// host_common.h -- host-side helpers of libpacmap_mi355x (errors, RAII, threads).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pacmap.h"

namespace pm {

// Internal failure carrying a pacmap_code; converted to pacmap_status at the
// C boundary (no exception crosses it).
struct PmError : std::runtime_error {
  int code;
  int gpu_error = 0;
  uint64_t expected[2] = {0, 0};
  uint64_t actual[2] = {0, 0};
  PmError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw PmError(code, msg); }

[[noreturn]] inline void fail_gpu(hipError_t e, const char *what, const char *file, int line) {
  PmError err(PACMAP_ERR_GPU, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" +
                                  std::to_string(line) + ")");
  err.gpu_error = static_cast<int>(e);
  throw err;
}

#define PM_HIP(expr)                                                 \
  do {                                                               \
    hipError_t pm_e_ = (expr);                                       \
    if (pm_e_ != hipSuccess) ::pm::fail_gpu(pm_e_, #expr, __FILE__, __LINE__); \
  } while (0)

// Per-thread cache of device blocks (pool.cpp).  A released block is kept
// under its size class and handed out again by a later allocation of the
// same thread instead of a hipMalloc / hipFree pair (hipFree synchronises the
// device).  Reuse is safe because all work of an entry point is ordered on
// its one stream and every entry point synchronises that stream before it
// returns; blocks never move between threads.
size_t dev_size_class(size_t bytes);
void *dev_acquire(size_t cls);        // cls from dev_size_class
void dev_recycle(void *p, size_t cls);
void dev_trim();                      // frees this thread's cached blocks
void graph_cache_reset();             // drops this thread's cached optimiser graph

// Buffers of the optimiser graph keep their own block from call to call
// (a slot per role), so identical calls launch identical kernel arguments
// and the graph can be replayed.  A slot block of the wrong size class goes
// back to the general cache.
enum Slot : int {
  kSlotNone = -1,
  kSlotCsrRow,
  kSlotCsrEnt,
  kSlotY,
  kSlotM,
  kSlotV,
  kSlotLossPart,
  kSlotSnap,
  kSlotTbEnt,
  kSlotTbOff,
  kSlotTbPart,
  kSlotSellPerm,
  kSlotSellOff,
  kSlotSellEnt,
  kSlotPageOff,
  kSlotPageEnt,
  kSlotCount
};
void *dev_acquire_slot(int slot, size_t cls);
void dev_release_slot(int slot, void *p, size_t cls);

template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t count, int slot = kSlotNone) { alloc(count, slot); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count, int slot = kSlotNone) {
    release();
    n_ = count;
    slot_ = slot;
    if (count) {
      cls_ = dev_size_class(count * sizeof(T));
      p_ = static_cast<T *>(slot >= 0 ? dev_acquire_slot(slot, cls_) : dev_acquire(cls_));
    }
  }
  void release() {
    if (p_) {
      if (slot_ >= 0)
        dev_release_slot(slot_, p_, cls_);
      else
        dev_recycle(p_, cls_);
    }
    p_ = nullptr;
    n_ = 0;
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T *p_ = nullptr;
  size_t n_ = 0;
  size_t cls_ = 0;
  int slot_ = kSlotNone;
};

// Thread-exit destructors release the thread's HIP resources unless the
// process is already exiting (see pool.cpp); note_runtime_use registers the
// hook that tells them.
void note_runtime_use();
bool process_exiting();

// The calling thread's stream: one non-blocking stream per thread, created
// on first use and destroyed when the thread exits (creating and destroying
// one per call costs tens of microseconds).  Every entry point orders all of
// its work on it and synchronises it before returning, so work left by a call
// that failed still runs before the next call's.
struct ThreadStream {
  hipStream_t s = nullptr;
  ~ThreadStream() {
    if (s && !process_exiting()) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};
class Stream {
 public:
  Stream() {
    static thread_local ThreadStream thread_stream;
    if (!thread_stream.s) {
      note_runtime_use();
      PM_HIP(hipStreamCreateWithFlags(&thread_stream.s, hipStreamNonBlocking));
    }
    s_ = thread_stream.s;
  }
  Stream(const Stream &) = delete;
  Stream &operator=(const Stream &) = delete;
  hipStream_t get() const { return s_; }
  void sync() const { PM_HIP(hipStreamSynchronize(s_)); }

 private:
  hipStream_t s_ = nullptr;
};

// A second per-thread stream for host-to-device copies that overlap the
// work of the calling thread's stream (same lifetime rules).
class CopyStream {
 public:
  CopyStream() {
    static thread_local ThreadStream thread_stream;
    if (!thread_stream.s) {
      note_runtime_use();
      PM_HIP(hipStreamCreateWithFlags(&thread_stream.s, hipStreamNonBlocking));
    }
    s_ = thread_stream.s;
  }
  CopyStream(const CopyStream &) = delete;
  CopyStream &operator=(const CopyStream &) = delete;
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

// Event pair timing a region of one stream.
// Work settled when an entry point returns (settle_pending): phase timings,
// and device flags copied into pinned host words.  Reading either as soon as
// its work is queued would make the host wait for the stream, and the GPU
// would then idle until the next launch reaches it.
struct PendingTime {
  hipEvent_t a, b;
  double *dst;
};
struct PendingFlag {
  hipEvent_t done;  // recorded after the copy into *word
  uint32_t *word;   // pinned host memory
  int code;
  const char *msg;
};
inline std::vector<PendingTime> &pending_times() {
  static thread_local std::vector<PendingTime> v;
  return v;
}
inline std::vector<PendingFlag> &pending_flags() {
  static thread_local std::vector<PendingFlag> v;
  return v;
}
constexpr int kPinnedWords = 64;
// kPinnedWords words of pinned host memory per thread, freed when the thread
// exits
struct PinnedWords {
  uint32_t *p = nullptr;
  ~PinnedWords() {
    if (p && !process_exiting()) (void)hipHostFree(p);
  }
};
inline uint32_t *pinned_words() {
  static thread_local PinnedWords w;
  if (!w.p) {
    note_runtime_use();
    PM_HIP(hipHostMalloc(reinterpret_cast<void **>(&w.p), kPinnedWords * sizeof(uint32_t), 0));
  }
  return w.p;
}
// Copies the device word `flag` into pinned memory on stream s; when the
// entry point returns, a nonzero value fails it with (code, msg).
inline void defer_flag_check(const uint32_t *flag, int code, const char *msg, hipStream_t s) {
  auto &v = pending_flags();
  if (v.size() >= static_cast<size_t>(kPinnedWords)) {  // no free word: check now
    uint32_t h = 0;
    PM_HIP(hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    if (h) fail(code, msg);
    return;
  }
  uint32_t *word = pinned_words() + v.size();
  *word = 0;
  hipEvent_t ev = nullptr;
  PM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  v.push_back({ev, word, code, msg});
  PM_HIP(hipMemcpyAsync(word, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PM_HIP(hipEventRecord(ev, s));
}
// Adds the pending timings to their destinations and checks the pending
// flags (read = true), or drops both; either way the lists end empty.
inline void settle_pending(bool read) {
  auto &t = pending_times();
  auto &f = pending_flags();
  int code = 0;
  const char *msg = nullptr;
  for (PendingTime &p : t) {
    float v = 0.f;
    if (read && hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&v, p.a, p.b) == hipSuccess)
      *p.dst += v;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (PendingFlag &p : f) {
    if (read && !code && hipEventSynchronize(p.done) == hipSuccess && *p.word) {
      code = p.code;
      msg = p.msg;
    }
    (void)hipEventDestroy(p.done);
  }
  t.clear();
  f.clear();
  if (code) fail(code, msg);
}

class Timer {
 public:
  Timer() {
    PM_HIP(hipEventCreate(&a_));
    PM_HIP(hipEventCreate(&b_));
  }
  ~Timer() {
    if (a_) (void)hipEventDestroy(a_);
    if (b_) (void)hipEventDestroy(b_);
  }
  Timer(const Timer &) = delete;
  Timer &operator=(const Timer &) = delete;
  void start(hipStream_t s) { PM_HIP(hipEventRecord(a_, s)); }
  void stop(hipStream_t s) { PM_HIP(hipEventRecord(b_, s)); }
  // waits for the stop event
  double ms() {
    PM_HIP(hipEventSynchronize(b_));
    float v = 0.f;
    PM_HIP(hipEventElapsedTime(&v, a_, b_));
    return v;
  }
  // adds the elapsed time to dst when the entry point returns; the timer
  // gives its events away and can no longer be used
  void add_to(double &dst) {
    pending_times().push_back({a_, b_, &dst});
    a_ = b_ = nullptr;
  }

 private:
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

inline void set_status(pacmap_status *st, int code, const char *msg) {
  if (!st) return;
  std::memset(st, 0, sizeof(*st));
  st->code = code;
  if (msg) {
    std::strncpy(st->message, msg, sizeof(st->message) - 1);
  }
}

// Runs body(); converts a PmError (or another exception) into the status
// code and message of the C boundary.
template <class F>
int guarded(pacmap_status *st, F &&body) {
  try {
    body();
    settle_pending(true);
    set_status(st, PACMAP_OK, nullptr);
    return PACMAP_OK;
  } catch (const PmError &e) {
    settle_pending(false);
    if (st) {
      set_status(st, e.code, e.what());
      st->gpu_error = e.gpu_error;
      st->expected[0] = e.expected[0];
      st->expected[1] = e.expected[1];
      st->actual[0] = e.actual[0];
      st->actual[1] = e.actual[1];
    }
    return e.code;
  } catch (const std::bad_alloc &) {
    settle_pending(false);
    set_status(st, PACMAP_ERR_GPU, "host allocation failed");
    return PACMAP_ERR_GPU;
  } catch (const std::exception &e) {
    settle_pending(false);
    set_status(st, PACMAP_ERR_INVALID_ARGUMENT, e.what());
    return PACMAP_ERR_INVALID_ARGUMENT;
  }
}

// Pointer kinds at the C boundary: a device allocation passed where host
// memory is expected (or host memory where device memory is) is reported as
// an invalid argument instead of faulting in a copy or a kernel.  Memory the
// runtime does not know (plain pageable host memory) counts as host memory;
// the error the query leaves for it is cleared, so it cannot surface in a
// later launch check.
inline bool device_pointer(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}
inline void require_host(const void *p, const char *what) {
  if (device_pointer(p)) fail(PACMAP_ERR_INVALID_ARGUMENT, std::string(what) + " is device memory; host memory expected");
}
inline void require_device(const void *p, const char *what) {
  if (p && !device_pointer(p)) fail(PACMAP_ERR_INVALID_ARGUMENT, std::string(what) + " is not device memory");
}

// Host-to-device copy on stream s: large copies go through the calling
// thread's pinned staging buffers (pool.cpp); src may be reused when this
// returns.  staging_release frees the thread's buffers.
void upload_bytes(void *dst, const void *src, size_t bytes, hipStream_t s);
// Uploads rows x row_bytes bytes in pieces of piece_rows whole rows (at most
// 32 MiB) through the pinned staging buffers, the DMA on copy stream cs
// (which first waits for the work queued on s).  After piece k is queued,
// on_piece(k, landed) is called with an event recorded on cs after it; wait
// on it (hipStreamWaitEvent) before the call returns to order work on the
// piece.
void upload_rows_pieces(void *dst, const void *src, uint64_t rows, size_t row_bytes, uint64_t piece_rows,
                        hipStream_t s, hipStream_t cs, const std::function<void(uint32_t, hipEvent_t)> &on_piece);
void staging_release();
// frees the calling thread's cached encoder transform weights
void mlp_transform_cache_release();
template <class T>
void upload(DevBuf<T> &dst, const T *src, size_t count, hipStream_t s) {
  if (count) upload_bytes(dst.get(), src, count * sizeof(T), s);
}
template <class T>
void download(T *dst, const T *src, size_t count, hipStream_t s) {
  if (count) PM_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

int host_threads();
// Runs fn(begin, end) over [0, total) split into contiguous ranges.
void parallel_for(uint64_t total, const std::function<void(uint64_t, uint64_t)> &fn,
                  int threads = 0);
double now_ms();

// --- reference arithmetic restated for the host (product copies) ----------
// weights.rs:52-81
void find_weights(float w_mn_init, uint64_t itr, uint64_t p1, uint64_t p2, float out[3]);
// adam.rs:46-47 (powi lowered to binary square-and-multiply)
float adam_lr_t(float lr, float beta1, float beta2, uint64_t itr);
// lib.rs:565-597 (throws on invalid counts)
void decide_num_pairs(uint64_t n, bool has_override, uint64_t override_k, float mn_ratio,
                      float fp_ratio, uint64_t out[3]);

// --- rand 0.8.5 SmallRng restatement (sampling.cpp) ------------------------
struct SmallRng {
  uint64_t s[4];
  static SmallRng seed_from_u64(uint64_t state);
  uint64_t next_u64();
  uint32_t next_u32() { return static_cast<uint32_t>(next_u64() >> 32); }
  uint32_t gen_range(uint32_t low, uint32_t high);
  double next_f64();  // [0, 1) with 53 bits
  double standard_normal();  // rand_distr 0.4.3 StandardNormal (ziggurat)
};

// A pair array (rows x 2 u32) in host memory.
struct PairList {
  const uint32_t *p = nullptr;
  uint64_t rows = 0;
};

}  // namespace pm
