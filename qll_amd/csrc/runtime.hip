// Runtime: device selection, the library stream, allocation, errors, lattice
// geometry and the deterministic-reduction scratch.
#include "qamd_internal.h"
#include "qamd_objects.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

namespace qamd {

static char g_err[1024] = "no error";
static hipStream_t g_stream = nullptr;
static int g_init = 0;
static double *g_red_base = nullptr;   // NSETS partial-sum sets
static int g_red_mb = 0;               // blocks per set

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  if (getenv("QAMD_VERBOSE")) fprintf(stderr, "qll_amd: %s\n", g_err);
}

int ensure_init() {
  if (g_init) return 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    set_error("no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
    return -1;
  }
  QAMD_HIP(hipSetDevice(0));
  // default priority stream, non-blocking w.r.t. the null stream
  QAMD_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_init = 1;
  return 0;
}

hipStream_t stream() { return g_stream; }

hipGraphExec_t capture_stream(int (*enqueue)(void *), void *ctx) {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  hipError_t e = hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal);
  int rc = 0;
  if (e == hipSuccess) {
    rc = enqueue(ctx);
    e = hipStreamEndCapture(g_stream, &g);   // ends the capture in every case
    const char *inj = getenv("QAMD_TEST_FAIL_CAPTURE");
    if (e == hipSuccess && inj && inj[0] == '1') e = hipErrorInvalidValue;
    if (e == hipSuccess && !rc) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  }
  if (g) (void)hipGraphDestroy(g);   // the executable graph keeps what it needs
  if (e == hipSuccess && !rc) return x;
  if (e != hipSuccess) set_error("graph capture: %s", hipGetErrorString(e));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(g_stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
    hipGraph_t g2 = nullptr;
    (void)hipStreamEndCapture(g_stream, &g2);
    if (g2) (void)hipGraphDestroy(g2);
  }
  if (x) (void)hipGraphExecDestroy(x);
  (void)hipGetLastError();
  return nullptr;
}
bool initialized() { return g_init != 0; }

bool slab_order(const Geom &g, SlabOrder &s, int run) {
  const int64_t plane = (int64_t)g.L[0] * g.L[1] / 2;   // sites of one parity per (z, t)
  if (plane % run != 0 || g.vh % run != 0) return false;
  const int P = (int)(plane / run);
  // pieces as thick as possible in both y and z (the gathers reach 3 sites)
  int best = -1;
  for (int py = 1; py <= NXCD; py *= 2) {
    const int pz = NXCD / py;
    if (P % py != 0 || g.L[2] % pz != 0) continue;
    const int ty = g.L[1] / py, tz = g.L[2] / pz, th = ty < tz ? ty : tz;
    if (th > best) {
      best = th;
      s.P = P;
      s.py = py;
      s.bpy = P / py;
      s.zs = tz;
      s.Lz = g.L[2];
      s.ny = s.nz = 1;
      s.Lt = g.L[3];
    }
  }
  return best > 0;
}

bool slab_split(SlabOrder &s, int ny, int nz) {
  if (ny < 1 || nz < 1 || s.bpy % ny != 0 || s.zs % nz != 0) return false;
  s.ny = ny;
  s.nz = nz;
  return true;
}

// Device blocks of 8 MB and more are kept for reuse when freed, up to 48 GB
// in all: once device memory has been cycled through, a hipMalloc of a
// multi-GB work field has taken 0.6-0.9 s on MI355X, and the solvers
// allocate their work fields per call (the 16-rhs Wilson CGNE at 32^3x64
// about 25 GB per solve: 0.23 s solves took 0.4-1.0 s).  A cached block is
// handed out again only for an allocation of exactly its size.  Freeing one
// first waits for the library stream, as hipFree did, so no queued kernel
// still uses it when it is handed out again.  A failed hipMalloc, and
// qamd_release_cache, free the cached blocks.  (The 8 MB floor takes in the
// single-rhs solvers' 12-25 MB vectors at 32^4: mixed Q16 CG ~1 % faster.)
constexpr size_t CACHE_MIN = (size_t)8 << 20;
constexpr size_t CACHE_MAX = (size_t)48 << 30;
static std::mutex g_cache_mu;
static std::unordered_map<void *, size_t> g_big;   // live blocks of CACHE_MIN or more
static std::multimap<size_t, void *> g_cache;      // freed blocks by size
static size_t g_cache_bytes = 0;

static void release_cache_locked() {
  for (auto &kv : g_cache) (void)hipFree(kv.second);
  g_cache.clear();
  g_cache_bytes = 0;
}

void *dev_alloc(size_t bytes) {
  if (ensure_init()) return nullptr;
  void *p = nullptr;
  if (bytes == 0) bytes = 256;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (bytes >= CACHE_MIN) {
    auto it = g_cache.find(bytes);
    if (it != g_cache.end()) {
      p = it->second;
      g_cache.erase(it);
      g_cache_bytes -= bytes;
      g_big[p] = bytes;
      return p;
    }
  }
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess && !g_cache.empty()) {   // make room and try once more
    (void)hipGetLastError();
    release_cache_locked();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  if (bytes >= CACHE_MIN) g_big[p] = bytes;
  return p;
}

void dev_free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_big.find(p);
  if (it != g_big.end()) {
    const size_t bytes = it->second;
    g_big.erase(it);
    if (g_cache_bytes + bytes <= CACHE_MAX && hipStreamSynchronize(g_stream) == hipSuccess) {
      g_cache.emplace(bytes, p);
      g_cache_bytes += bytes;
      return;
    }
  }
  (void)hipFree(p);
}

RedScratch red_set(int set) {
  RedScratch r;
  r.part = g_red_base + (size_t)set * MAXRED * g_red_mb;
  r.max_blocks = g_red_mb;
  return r;
}

int red_reserve(int nblocks) {
  if (ensure_init()) return -1;
  if (nblocks <= g_red_mb) return 0;
  // grow: wait for pending work that may still use the old scratch
  QAMD_HIP(hipStreamSynchronize(g_stream));
  dev_free(g_red_base);
  int mb = 1024;
  while (mb < nblocks) mb *= 2;
  g_red_base = (double *)dev_alloc(sizeof(double) * NSETS * MAXRED * mb);
  if (!g_red_base) {
    g_red_mb = 0;
    return -1;
  }
  g_red_mb = mb;
  return 0;
}

void geom_init(Geom &g, const int L[4]) {
  g.V = 1;
  for (int i = 0; i < 4; i++) {
    g.L[i] = L[i];
    g.V *= L[i];
  }
  g.vh = g.V / 2;
  g.vh_pad = (g.vh + WAVE - 1) / WAVE * WAVE;
  g.stride[0] = 1;
  g.stride[1] = L[0];
  g.stride[2] = L[0] * L[1];
  g.stride[3] = L[0] * L[1] * L[2];
  g.dxh.init(L[0] / 2);
  g.dy.init(L[1]);
  g.dz.init(L[2]);
}

}  // namespace qamd

using namespace qamd;

extern "C" {

const char *qamd_last_error(void) { return g_err; }

int qamd_init(void) { return ensure_init(); }

int qamd_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (g_stream) (void)hipStreamSynchronize(g_stream);
  release_cache_locked();
  return 0;
}

int qamd_sync(void) {
  if (ensure_init()) return -1;
  QAMD_HIP(hipStreamSynchronize(g_stream));
  return 0;
}

int qamd_device_name(char *buf, int len) {
  if (ensure_init()) return -1;
  hipDeviceProp_t p;
  QAMD_HIP(hipGetDeviceProperties(&p, 0));
  snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return 0;
}

void *qamd_malloc(size_t bytes) { return dev_alloc(bytes); }
void qamd_free(void *p) { dev_free(p); }

int qamd_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  if (ensure_init() || h2d_pageable(dst, src, bytes)) return -1;
  QAMD_HIP(hipStreamSynchronize(g_stream));
  return 0;
}

int qamd_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if (ensure_init()) return -1;
  return d2h_pageable(dst, src, bytes);
}

qamd_lattice *qamd_lattice_create(const int dims[4]) {
  for (int i = 0; i < 4; i++) {
    if (dims[i] < 4 || (dims[i] & 1)) {
      set_error("lattice extent %d = %d: extents must be even and >= 4", i, dims[i]);
      return nullptr;
    }
  }
  int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  if (V / 2 >= ((int64_t)1 << 30)) {
    set_error("lattice of %lld sites exceeds the 2^31 site limit", (long long)V);
    return nullptr;
  }
  if (ensure_init()) return nullptr;
  qamd_lattice *lat = (qamd_lattice *)calloc(1, sizeof(qamd_lattice));
  geom_init(lat->g, dims);
  return lat;
}

void qamd_lattice_destroy(qamd_lattice *lat) { free(lat); }

int64_t qamd_lattice_sites(const qamd_lattice *lat) { return lat ? lat->g.V : 0; }

}  // extern "C"

// ------------------------------------------------------------- autotuner
// Launch-shape choices measured on first use and cached per key (lattice,
// precisions, block size, ...): every candidate is launched once to warm up,
// then timed over three launches with events on the library stream, and the
// fastest is kept for the rest of the process.  The kernels tuned this way
// are out of place, so repeating a launch leaves the same result.  No tuning
// while the stream is being captured into a graph (the cached or default
// choice is used) or when QAMD_TUNE=0.
namespace qamd {
static std::map<std::string, int> g_tuned;
static int g_tune_on = -1;
static int g_tune_force = -1;   // >= 0: every tuned launch uses this candidate

static bool tuning_enabled() {
  if (g_tune_on < 0) {
    const char *e = getenv("QAMD_TUNE");
    g_tune_on = !(e && e[0] == '0');
  }
  return g_tune_on;
}

int tune_pick(const char *key, int ncand, int def, int (*run)(int, void *), void *ctx) {
  if (g_tune_force >= 0) return g_tune_force < ncand ? g_tune_force : ncand - 1;
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) return it->second;
  if (ncand <= 1 || !tuning_enabled()) return def;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(g_stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return def;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return def;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return def;
  }
  int best = def;
  float tbest = 1e30f;
  for (int c = 0; c < ncand; c++) {
    if (run(c, ctx)) continue;
    (void)hipEventRecord(e0, g_stream);
    int bad = 0;
    for (int k = 0; k < 3; k++) bad |= run(c, ctx);
    (void)hipEventRecord(e1, g_stream);
    if (bad || hipEventSynchronize(e1) != hipSuccess) continue;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < tbest) {
      tbest = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_tuned[key] = best;
  if (getenv("QAMD_VERBOSE")) fprintf(stderr, "qll_amd tune %s: candidate %d (%.4f ms)\n", key, best, tbest / 3);
  return best;
}
}  // namespace qamd

extern "C" {
int qamd_set_tuning(int mode) {
  qamd::g_tune_on = mode ? 1 : 0;
  qamd::g_tune_force = mode >= 2 ? mode - 2 : -1;
  if (mode != 1) qamd::g_tuned.clear();
  return 0;
}
}
