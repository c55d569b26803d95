// Host <-> device transfers of caller arrays in pageable host memory (the QLL
// API's fields, lexicographic and site-table imports): a persistent device
// staging buffer and two pinned host buffers with a pool of copy threads.
// A transfer is cut into chunks; for an upload, threads copy chunk k into
// pinned buffer k%2 while the DMA engine moves chunk k-1 to the device, and
// downloads run the same pipeline backwards.  Nothing is allocated per call
// once the buffers have reached the largest transfer size.
#include "qamd_internal.h"

#include <condition_variable>
#include <mutex>
#include <stdlib.h>
#include <string.h>
#include <thread>
#include <vector>

namespace qamd {

namespace {

constexpr size_t CHUNK = (size_t)16 << 20;   // bytes per pinned buffer
constexpr size_t SMALL = (size_t)256 << 10;  // below this one thread copies

// Workers copy slices 1..n of a memcpy while the caller copies slice 0.  The
// pool lives until the process ends (threads are detached and idle on a
// condition variable between copies).
class CopyPool {
 public:
  explicit CopyPool(int nthreads) : nw_(nthreads) {
    for (int i = 0; i < nthreads; i++) std::thread(&CopyPool::work, this, i + 1).detach();
  }
  void copy(void *dst, const void *src, size_t n) {
    if (nw_ == 0 || n < SMALL) {
      memcpy(dst, src, n);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = (char *)dst;
      src_ = (const char *)src;
      n_ = n;
      pending_ = nw_;
      gen_++;
    }
    go_.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  void part(int k) {   // slice k of nw_ + 1, boundaries on 64-byte multiples
    const size_t np = (size_t)nw_ + 1;
    size_t a = (n_ * k / np) & ~(size_t)63, b = k + 1 == (int)np ? n_ : (n_ * (k + 1) / np) & ~(size_t)63;
    if (b > a) memcpy(dst_ + a, src_ + a, b - a);
  }
  void work(int k) {
    unsigned long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        go_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      part(k);
      std::lock_guard<std::mutex> lk(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex m_;
  std::condition_variable go_, done_;
  unsigned long gen_ = 0;
  int pending_ = 0, nw_ = 0;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t n_ = 0;
};

CopyPool *g_pool = nullptr;
void *g_pin[2] = {nullptr, nullptr};
hipEvent_t g_ev[2];
void *g_stage = nullptr;
size_t g_stage_bytes = 0;

int copy_threads() {
  const char *s = getenv("QAMD_COPY_THREADS");
  int n = s ? atoi(s) : (int)std::thread::hardware_concurrency();
  if (n > 8) n = 8;   // host memcpy bandwidth saturates at a few threads
  return n < 1 ? 1 : n;
}

int pinned_init() {
  if (g_pin[0]) return 0;
  if (ensure_init()) return -1;
  for (int b = 0; b < 2; b++) {
    QAMD_HIP(hipHostMalloc(&g_pin[b], CHUNK, hipHostMallocDefault));
    QAMD_HIP(hipEventCreateWithFlags(&g_ev[b], hipEventDisableTiming));
  }
  g_pool = new CopyPool(copy_threads() - 1);
  return 0;
}

}  // namespace

void *stage_dev(size_t bytes) {
  if (ensure_init()) return nullptr;
  if (bytes <= g_stage_bytes) return g_stage;
  // grow: pending work on the stream may still read the old buffer
  if (hipStreamSynchronize(stream()) != hipSuccess) {
    set_error("stage_dev: stream synchronisation failed");
    return nullptr;
  }
  dev_free(g_stage);
  g_stage_bytes = 0;
  size_t nb = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
  g_stage = dev_alloc(nb);
  if (g_stage) g_stage_bytes = nb;
  return g_stage;
}

int h2d_pageable(void *dst, const void *src, size_t bytes) {
  if (pinned_init()) return -1;
  const char *s = (const char *)src;
  char *d = (char *)dst;
  for (size_t off = 0, k = 0; off < bytes; off += CHUNK, k++) {
    const int b = (int)(k & 1);
    const size_t n = bytes - off < CHUNK ? bytes - off : CHUNK;
    QAMD_HIP(hipEventSynchronize(g_ev[b]));   // the DMA that last read this buffer
    g_pool->copy(g_pin[b], s + off, n);
    QAMD_HIP(hipMemcpyAsync(d + off, g_pin[b], n, hipMemcpyHostToDevice, stream()));
    QAMD_HIP(hipEventRecord(g_ev[b], stream()));
  }
  return 0;
}

int d2h_pageable(void *dst, const void *src, size_t bytes) {
  if (pinned_init()) return -1;
  const char *s = (const char *)src;
  char *d = (char *)dst;
  size_t poff = 0, pn = 0;
  int pb = -1;
  for (size_t off = 0, k = 0; off < bytes; off += CHUNK, k++) {
    const int b = (int)(k & 1);
    const size_t n = bytes - off < CHUNK ? bytes - off : CHUNK;
    QAMD_HIP(hipEventSynchronize(g_ev[b]));   // an upload may still read this buffer
    QAMD_HIP(hipMemcpyAsync(g_pin[b], s + off, n, hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipEventRecord(g_ev[b], stream()));
    if (pb >= 0) {
      QAMD_HIP(hipEventSynchronize(g_ev[pb]));
      g_pool->copy(d + poff, g_pin[pb], pn);
    }
    pb = b;
    poff = off;
    pn = n;
  }
  if (pb >= 0) {
    QAMD_HIP(hipEventSynchronize(g_ev[pb]));
    g_pool->copy(d + poff, g_pin[pb], pn);
  }
  return 0;
}

}  // namespace qamd
