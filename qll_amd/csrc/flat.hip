// Host-layout helpers behind the QLL drop-in API: a generic QLL-layout
// shift/transport kernel, FP64-accumulated sums of host arrays, and flat
// device BLAS-1 for solvers that call host Linop callbacks.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_site.h"
#include "qamd_kernels.h"

#include <string.h>

namespace qamd {

template <class H>
__global__ void k_qll_shift(QllLay lay, int disp0, int disp1, int disp2, int disp3, int nelem, int nc, int td,
                            const H *u, const H *src, H *dest, int b0, int b1, int acc, int adj) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t V = (int64_t)lay.phys[0] * lay.phys[1] * lay.phys[2] * lay.phys[3];
  if (s >= V) return;
  int c[4];
  int64_t t = s;
  for (int i = 0; i < 4; i++) {
    c[i] = (int)(t % lay.phys[i]);
    t /= lay.phys[i];
  }
  int oi, ii;
  qll_index(lay, c, oi, ii);
  if (oi < b0 || oi >= b1) return;
  const int disp[4] = {disp0, disp1, disp2, disp3};
  int cs[4];
  for (int i = 0; i < 4; i++) cs[i] = ((c[i] - disp[i]) % lay.phys[i] + lay.phys[i]) % lay.phys[i];
  int so, si;
  qll_index(lay, cs, so, si);
  const int NI = lay.NI;
  const int64_t db = (int64_t)nelem * NI * oi + ii, sb = (int64_t)nelem * NI * so + si;
  double r[72];
  if (nelem > 72) return;
  if (u && td > 0) {
    // r[i,j,k] = sum_l m[j,l] v[i,l,k]  (complex pairs, tdv = 2)
    const int64_t mb = (int64_t)(2 * nc * nc) * NI * oi + ii;
    int ldv = nelem / (nc * td);
    for (int i = 0; i < ldv; i++)
      for (int j = 0; j < nc; j++) {
        double re = 0, im = 0;
        for (int l = 0; l < nc; l++) {
          double mr = (double)u[mb + (int64_t)NI * 2 * (j + l * nc)];
          double mi = (double)u[mb + (int64_t)NI * (2 * (j + l * nc) + 1)];
          double vr = (double)src[sb + (int64_t)NI * ((i * nc + l) * td)];
          double vi = (double)src[sb + (int64_t)NI * ((i * nc + l) * td + 1)];
          re += mr * vr - mi * vi;
          im += mr * vi + mi * vr;
        }
        r[(i * nc + j) * td] = re;
        r[(i * nc + j) * td + 1] = im;
      }
  } else {
    for (int e = 0; e < nelem; e++) r[e] = (double)src[sb + (int64_t)NI * e];
  }
  if (adj && nelem == 2 * nc * nc) {
    double a[72];
    for (int rr = 0; rr < nc; rr++)
      for (int cc = 0; cc < nc; cc++) {
        a[2 * (rr + cc * nc)] = r[2 * (cc + rr * nc)];
        a[2 * (rr + cc * nc) + 1] = -r[2 * (cc + rr * nc) + 1];
      }
    for (int e = 0; e < nelem; e++) r[e] = a[e];
  }
  for (int e = 0; e < nelem; e++) {
    int64_t k = db + (int64_t)NI * e;
    dest[k] = (H)(acc ? (double)dest[k] + r[e] : r[e]);
  }
}

struct FinOne {
  double *dst;
  __device__ void operator()(const double *t) const { *dst = t[0]; }
};

template <class T>
__global__ void k_flat_sum(const T *x, const T *y, int64_t n, int mode, RedScratch rs) {
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double a = (double)x[i];
    acc[0] += mode == 0 ? a : a * (double)(y ? y[i] : x[i]);
  }
  grid_partial<1>(acc, rs);
}

template <class T>
__global__ void k_flat_axpby(int64_t n, double a, const T *x, double b, T *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)(b == 0.0 ? a * (double)x[i] : a * (double)x[i] + b * (double)y[i]);
}

}  // namespace qamd

using namespace qamd;

extern "C" {

int qamd_qll_shift(const qamd_qll_layout *l, const int disp[4], int nelem, int nc, int td, const void *u,
                   const void *src, void *dest, int hprec, int beginOuter, int endOuter, int accumulate,
                   int adjoint) {
  if (ensure_init()) return -1;
  if (nelem < 1 || nelem > 72) { set_error("qll_shift: nelem %d out of range", nelem); return -1; }
  if (u && td > 0 && (td != 2 || nc < 1 || nelem % (nc * td))) {
    set_error("qll_shift: transport needs td = 2 and nelem a multiple of 2*nc");
    return -1;
  }
  QllLay lay = to_lay(l);
  int64_t V = (int64_t)l->physGeom[0] * l->physGeom[1] * l->physGeom[2] * l->physGeom[3];
  size_t esz = hprec == QAMD_F64 ? 8 : 4;
  size_t bytes = (size_t)V * nelem * esz, ubytes = (size_t)V * 2 * nc * nc * esz;
  const bool mat = u && td > 0;
  // one device buffer for source, destination and matrices (grown as needed)
  size_t off_d = (bytes + 255) & ~(size_t)255, off_u = 2 * off_d;
  char *buf = (char *)stage_dev(off_u + (mat ? ubytes : 0));
  if (!buf) return -1;
  void *ds = buf, *dd = buf + off_d, *du = mat ? buf + off_u : nullptr;
  if (h2d_pageable(ds, src, bytes) || h2d_pageable(dd, dest, bytes) || (mat && h2d_pageable(du, u, ubytes)))
    return -1;
  unsigned nb = (unsigned)((V + 255) / 256);
  if (hprec == QAMD_F64)
    hipLaunchKernelGGL(k_qll_shift<double>, dim3(nb), dim3(256), 0, stream(), lay, disp[0], disp[1], disp[2],
                       disp[3], nelem, nc, td, (const double *)du, (const double *)ds, (double *)dd, beginOuter,
                       endOuter, accumulate, adjoint);
  else
    hipLaunchKernelGGL(k_qll_shift<float>, dim3(nb), dim3(256), 0, stream(), lay, disp[0], disp[1], disp[2],
                       disp[3], nelem, nc, td, (const float *)du, (const float *)ds, (float *)dd, beginOuter,
                       endOuter, accumulate, adjoint);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("qll_shift: %s", hipGetErrorString(e));
    return -1;
  }
  return d2h_pageable(dest, dd, bytes);
}

static int flat_reduce(int64_t n, int prec, const void *x, const void *y, int mode, double *out) {
  int nb = 1024;
  if (red_reserve(nb)) return -1;
  double *slot = scalar_slot();
  if (!slot) return -1;
  RedScratch rs = red_set(0);
  if (prec == QAMD_F64)
    hipLaunchKernelGGL(k_flat_sum<double>, dim3(nb), dim3(256), 0, stream(), (const double *)x, (const double *)y,
                       n, mode, rs);
  else
    hipLaunchKernelGGL(k_flat_sum<float>, dim3(nb), dim3(256), 0, stream(), (const float *)x, (const float *)y, n,
                       mode, rs);
  launch_finish<1>(rs, nb, FinOne{slot + 8}, stream());
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipMemcpyAsync(out, slot + 8, sizeof(double), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int qamd_host_sum(const void *host, int64_t n, int hprec, int squares, double *out) {
  if (ensure_init()) return -1;
  size_t bytes = (size_t)n * (hprec == QAMD_F64 ? 8 : 4);
  void *d = stage_dev(bytes);
  if (!d || h2d_pageable(d, host, bytes)) return -1;
  return flat_reduce(n, hprec, d, nullptr, squares ? 1 : 0, out);
}

int qamd_flat_axpby(int64_t n, int prec, double a, const void *x, double b, void *y) {
  if (ensure_init()) return -1;
  unsigned nb = (unsigned)((n + 255) / 256);
  if (nb > 65536) nb = 65536;
  if (nb == 0) return 0;
  if (prec == QAMD_F64)
    hipLaunchKernelGGL(k_flat_axpby<double>, dim3(nb), dim3(256), 0, stream(), n, a, (const double *)x, b,
                       (double *)y);
  else
    hipLaunchKernelGGL(k_flat_axpby<float>, dim3(nb), dim3(256), 0, stream(), n, a, (const float *)x, b,
                       (float *)y);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int qamd_flat_dot(int64_t n, int prec, const void *x, const void *y, double *out) {
  if (ensure_init()) return -1;
  return flat_reduce(n, prec, x, y, 1, out);
}

}  // extern "C"
