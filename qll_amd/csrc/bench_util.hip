// Measurement helpers: a plain streaming copy kernel of known size (the
// STREAM-copy reference for the roofline, and the calibration kernel for
// rocprof's FETCH_SIZE / WRITE_SIZE counters).
#include "qamd_internal.h"

#include <chrono>
#include <stdlib.h>
#include <string.h>

namespace qamd {

// one 16-byte element per thread, non-temporal load (measured on MI355X: one
// element per thread with 8192+ blocks streams faster than 2-8 elements per
// thread: 6.5 against 5.8-5.9 TB/s read + write)
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = ntload(a + i);
}

}  // namespace qamd

using namespace qamd;

static double wall_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

// copy 'bytes' (rounded down to 16 B) nrep times; returns the copy bandwidth
// counting read + write bytes, in GB/s, and the mean ms per copy
int qamd_stream_copy(size_t bytes, int nrep, double *gbs, double *ms_per_copy) {
  if (ensure_init()) return -1;
  int64_t n = (int64_t)(bytes / 16);
  double2 *a = (double2 *)dev_alloc(n * 16), *b = (double2 *)dev_alloc(n * 16);
  if (!a || !b) {
    dev_free(a);
    dev_free(b);
    return -1;
  }
  QAMD_HIP(hipMemsetAsync(a, 0, n * 16, stream()));
  hipEvent_t e0, e1;
  QAMD_HIP(hipEventCreate(&e0));
  QAMD_HIP(hipEventCreate(&e1));
  int nb = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, stream(), a, b, n);   // warm-up
  QAMD_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < nrep; i++) hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, stream(), a, b, n);
  QAMD_HIP(hipEventRecord(e1, stream()));
  QAMD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  double t = ms / nrep;
  if (ms_per_copy) *ms_per_copy = t;
  if (gbs) *gbs = 2.0 * n * 16 / (t * 1e-3) / 1e9;
  return 0;
}

// Host <-> device copy rates in GB/s: from pinned host memory
// (hipHostMalloc, the DMA engines alone) and from pageable memory through the
// library's staging pipeline (the path of every QLL-API transfer).
int qamd_host_copy_rate(size_t bytes, int nrep, double rates[4]) {
  if (ensure_init()) return -1;
  void *dv = dev_alloc(bytes), *hp = nullptr;
  char *hg = (char *)malloc(bytes);
  if (!dv || !hg || hipHostMalloc(&hp, bytes, hipHostMallocDefault) != hipSuccess) {
    dev_free(dv);
    free(hg);
    set_error("host_copy_rate: allocation of %zu bytes failed", bytes);
    return -1;
  }
  memset(hg, 1, bytes);
  memset(hp, 1, bytes);
  int rc = 0;
  for (int k = 0; k < 4 && !rc; k++) {   // 0 pinned h2d, 1 pinned d2h, 2 pageable h2d, 3 pageable d2h
    double t0 = 0;
    for (int i = -1; i < nrep && !rc; i++) {   // i = -1: warm-up
      if (i == 0) t0 = wall_s();
      hipError_t e = hipSuccess;
      if (k == 0) e = hipMemcpyAsync(dv, hp, bytes, hipMemcpyHostToDevice, stream());
      else if (k == 1) e = hipMemcpyAsync(hp, dv, bytes, hipMemcpyDeviceToHost, stream());
      else if (k == 2) rc = h2d_pageable(dv, hg, bytes);
      else rc = d2h_pageable(hg, dv, bytes);
      if (e == hipSuccess) e = hipStreamSynchronize(stream());
      if (e != hipSuccess) {
        set_error("host_copy_rate: %s", hipGetErrorString(e));
        rc = -1;
      }
    }
    rates[k] = (double)bytes * nrep / (wall_s() - t0) / 1e9;
  }
  dev_free(dv);
  (void)hipHostFree(hp);
  free(hg);
  return rc;
}

}  // extern "C"
