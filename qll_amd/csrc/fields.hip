// Fields: creation, host import/export (lexicographic and QLL layouts),
// Gaussian sources, FP64-accumulated norms/dots and axpby.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_site.h"
#include "qamd_kernels.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

namespace qamd {

size_t prec_bytes(int prec) {
  return prec == QAMD_F64 ? 16 : prec == QAMD_F32 ? 8 : 4;   // FP16 and Q16: 4 bytes
}

// the (site, rhs) of this thread in a (sites/256, nrhs) grid: rhs in
// blockIdx.y for colour vectors; for spinors, whose rhs are adjacent in
// memory, the flattened thread index with the rhs fastest
__device__ __forceinline__ bool site_rhs(const Field &f, int &cb, int &rhs) {
  if (!f.rhs_inner) {
    cb = blockIdx.x * blockDim.x + threadIdx.x;
    rhs = blockIdx.y;
  } else {
    const int64_t i = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    cb = (int)(i / f.nrhs);
    rhs = (int)(i - (int64_t)cb * f.nrhs);
  }
  return cb < f.g->vh;
}

template <class H, int NC>
__global__ void k_import(Field f, const H *src, SrcMap m) {
  int cb, rhs;
  if (!site_rhs(f, cb, rhs)) return;
  int64_t base, stride;
  m.map(*f.g, rhs, f.parity, cb, base, stride);
  double2 v[NC];
#pragma unroll
  for (int j = 0; j < NC; j++)
    v[j] = make_double2((double)src[base + (2 * j) * stride], (double)src[base + (2 * j + 1) * stride]);
  site_store<NC>(f, rhs, cb, v);
}

template <class H, int NC>
__global__ void k_export(Field f, H *dst, SrcMap m) {
  int cb, rhs;
  if (!site_rhs(f, cb, rhs)) return;
  int64_t base, stride;
  m.map(*f.g, rhs, f.parity, cb, base, stride);
  double2 v[NC];
  site_load<NC>(f, rhs, cb, v);
#pragma unroll
  for (int j = 0; j < NC; j++) {
    dst[base + (2 * j) * stride] = (H)v[j].x;
    dst[base + (2 * j + 1) * stride] = (H)v[j].y;
  }
}

template <int NC>
__global__ void k_gaussian(Field f, uint64_t seed, uint64_t stream) {
  int cb, rhs;
  if (!site_rhs(f, cb, rhs)) return;
  Site s = site_of(*f.g, cb, f.parity);
  double2 v[NC];
#pragma unroll
  for (int j = 0; j < NC; j++) {
    double a, b;
    gauss2(seed, stream + rhs, (uint64_t)s.lex, j, a, b);
    v[j] = make_double2(a * 0.70710678118654752440, b * 0.70710678118654752440);
  }
  site_store<NC>(f, rhs, cb, v);
}

// --------------------------------------------------------------- BLAS-1
template <int NC>
__global__ void k_norm_dot(Field a, Field b, int rhs, int dot, RedScratch rs) {
  double acc[1] = {0.0};
  for (int cb = blockIdx.x * blockDim.x + threadIdx.x; cb < a.g->vh; cb += gridDim.x * blockDim.x) {
    double2 va[NC], vb[NC];
    site_load<NC>(a, rhs, cb, va);
    if (dot) site_load<NC>(b, rhs, cb, vb);
#pragma unroll
    for (int j = 0; j < NC; j++) {
      if (dot) acc[0] += va[j].x * vb[j].x + va[j].y * vb[j].y;
      else acc[0] += va[j].x * va[j].x + va[j].y * va[j].y;
    }
  }
  grid_partial<1>(acc, rs);
}

struct StoreTo {
  double *dst;
  __device__ void operator()(const double *t) const { *dst = t[0]; }
};

template <int NC>
__global__ void k_axpby(double a, Field x, double b, Field y) {
  int cb, rhs;
  if (!site_rhs(y, cb, rhs)) return;
  double2 vx[NC], vy[NC];
  site_load<NC>(x, rhs, cb, vx);
  if (b != 0.0) site_load<NC>(y, rhs, cb, vy);
#pragma unroll
  for (int j = 0; j < NC; j++)
    vy[j] = (b != 0.0) ? make_double2(a * vx[j].x + b * vy[j].x, a * vx[j].y + b * vy[j].y)
                       : make_double2(a * vx[j].x, a * vx[j].y);
  site_store<NC>(y, rhs, cb, vy);
}

// ------------------------------------------------------------- helpers
inline dim3 site_grid(const Geom &g, int nrhs, int bs = 256) {
  return dim3((unsigned)((g.vh + bs - 1) / bs), (unsigned)nrhs);
}

double *scalar_slot() {
  static double *d = nullptr;
  if (!d) d = (double *)dev_alloc(64 * sizeof(double));
  return d;
}

int reduce_field(const Field &a, const Field *b, int rhs, double *out) {
  int nb = 1024;
  if (red_reserve(nb)) return -1;
  double *slot = scalar_slot();
  if (!slot) return -1;
  StoreTo fin{slot};
  Field bb = b ? *b : a;
  RedScratch rs = red_set(0);
  hipLaunchKernelGGL((a.ncomp == 3 ? k_norm_dot<3> : k_norm_dot<12>), dim3(nb), dim3(256), 0, stream(), a, bb, rhs,
                     b ? 1 : 0, rs);
  launch_finish<1>(rs, nb, fin, stream());
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipMemcpyAsync(out, slot, sizeof(double), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// need_parity: 0/1 = every site of that parity must appear, 2 = every site
int site_table_inverse(const qamd_lattice *lat, const int *coords, int64_t nsites, int need_parity, int **inv) {
  const Geom &g = lat->g;
  *inv = nullptr;
  if (!coords || nsites < 1 || nsites > g.V) {
    set_error("site table: %lld sites for a lattice of %lld", (long long)nsites, (long long)g.V);
    return -1;
  }
  int *h = (int *)malloc(sizeof(int) * g.V);
  if (!h) {
    set_error("site table: host allocation failed");
    return -1;
  }
  for (int64_t i = 0; i < g.V; i++) h[i] = -1;
  int64_t have[2] = {0, 0};
  for (int64_t i = 0; i < nsites; i++) {
    const int *c = coords + 4 * i;
    int64_t lex = 0;
    for (int d = 3; d >= 0; d--) {
      if (c[d] < 0 || c[d] >= g.L[d]) {
        set_error("site table: site %lld has coordinate %d = %d outside 0..%d", (long long)i, d, c[d], g.L[d] - 1);
        free(h);
        return -1;
      }
      lex = lex * g.L[d] + c[d];
    }
    if (h[lex] >= 0) {
      set_error("site table: sites %d and %lld have the same coordinates", h[lex], (long long)i);
      free(h);
      return -1;
    }
    h[lex] = (int)i;
    have[(c[0] + c[1] + c[2] + c[3]) & 1]++;
  }
  if ((need_parity != 1 && have[0] != g.vh) || (need_parity != 0 && have[1] != g.vh)) {
    set_error("site table: %lld even and %lld odd sites listed, the lattice has %lld of each",
              (long long)have[0], (long long)have[1], (long long)g.vh);
    free(h);
    return -1;
  }
  int *d = (int *)dev_alloc(sizeof(int) * g.V);
  hipError_t e = d ? hipMemcpy(d, h, sizeof(int) * g.V, hipMemcpyHostToDevice) : hipErrorOutOfMemory;
  free(h);
  if (e != hipSuccess) {
    dev_free(d);
    if (d) set_error("site table: copy failed: %s", hipGetErrorString(e));
    return -1;
  }
  *inv = d;
  return 0;
}

}  // namespace qamd

using namespace qamd;

namespace qamd {
const Geom *dev_geom(qamd_lattice *lat) {
  // one device copy per distinct geometry, kept until the process ends
  struct Ent { Geom h; Geom *d; };
  static Ent tab[256];
  static int n = 0;
  for (int i = 0; i < n; i++)
    if (memcmp(&tab[i].h, &lat->g, sizeof(Geom)) == 0) return tab[i].d;
  if (n >= 256) {
    set_error("more than 256 distinct lattice geometries");
    return nullptr;
  }
  Geom *d = (Geom *)dev_alloc(sizeof(Geom));
  if (!d) return nullptr;
  if (hipMemcpy(d, &lat->g, sizeof(Geom), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("dev_geom: copy failed");
    return nullptr;
  }
  tab[n].h = lat->g;
  tab[n].d = d;
  n++;
  return d;
}
}  // namespace qamd

extern "C" {

qamd_field *qamd_field_create(qamd_lattice *lat, int prec, int ncomp, int nrhs, int parity) {
  if (!lat) { set_error("field_create: NULL lattice"); return nullptr; }
  if (prec < QAMD_F64 || prec > QAMD_Q16) { set_error("field_create: bad precision %d", prec); return nullptr; }
  if (ncomp != 3 && ncomp != 12) { set_error("field_create: ncomp must be 3 or 12"); return nullptr; }
  if (nrhs < 1 || nrhs > 64) { set_error("field_create: nrhs must be 1..64"); return nullptr; }
  if (parity != 0 && parity != 1) { set_error("field_create: parity must be 0 or 1"); return nullptr; }
  const Geom *dg = dev_geom(lat);
  if (!dg) return nullptr;
  qamd_field *f = (qamd_field *)calloc(1, sizeof(qamd_field));
  f->lat = lat;
  f->f.prec = prec;
  f->f.ncomp = ncomp;
  f->f.nrhs = nrhs;
  f->f.parity = parity;
  f->f.g = dg;
  f->f.vp = lat->g.vh_pad;
  f->f.rhs_inner = (ncomp == 12);
  f->f.kc = prec == QAMD_F64 ? 1 : prec == QAMD_F32 ? 2 : 4;
  int64_t n = (int64_t)nrhs * ncomp * lat->g.vh_pad;
  f->f.data = dev_alloc(n * prec_bytes(prec));
  const bool normed = (prec == QAMD_F16 || prec == QAMD_Q16);
  if (normed) f->f.norm = (float *)dev_alloc(sizeof(float) * nrhs * lat->g.vh_pad);
  if (!f->f.data || (normed && !f->f.norm)) {
    dev_free(f->f.data);
    dev_free(f->f.norm);
    free(f);
    return nullptr;
  }
  if (qamd_field_zero(f)) {
    qamd_field_destroy(f);
    return nullptr;
  }
  return f;
}

void qamd_field_destroy(qamd_field *f) {
  if (!f) return;
  (void)hipStreamSynchronize(stream());
  dev_free(f->f.data);
  dev_free(f->f.norm);
  free(f);
}

int qamd_field_zero(qamd_field *f) {
  int64_t n = f->f.elems();
  QAMD_HIP(hipMemsetAsync(f->f.data, 0, n * prec_bytes(f->f.prec), stream()));
  if (f->f.norm)
    QAMD_HIP(hipMemsetAsync(f->f.norm, 0, sizeof(float) * f->f.nrhs * f->lat->g.vh_pad, stream()));
  return 0;
}

int qamd_field_copy(qamd_field *dst, const qamd_field *src) {
  if (dst->lat->g.V != src->lat->g.V || dst->f.ncomp != src->f.ncomp ||
      dst->f.nrhs != src->f.nrhs || dst->f.parity != src->f.parity) {
    set_error("field_copy: shape mismatch");
    return -1;
  }
  if (dst->f.prec == src->f.prec) {
    QAMD_HIP(hipMemcpyAsync(dst->f.data, src->f.data, src->f.elems() * prec_bytes(src->f.prec),
                            hipMemcpyDeviceToDevice, stream()));
    if (src->f.norm)
      QAMD_HIP(hipMemcpyAsync(dst->f.norm, src->f.norm, sizeof(float) * src->f.nrhs * src->lat->g.vh_pad,
                              hipMemcpyDeviceToDevice, stream()));
    return 0;
  }
  return qamd_field_axpby(1.0, src, 0.0, dst);
}

static int import_common(qamd_field *f, const void *host, size_t bytes, int hprec, SrcMap m) {
  void *stage = stage_dev(bytes);
  if (!stage || h2d_pageable(stage, host, bytes)) return -1;
  dim3 grid = site_grid(f->lat->g, f->f.nrhs);
  const bool c3 = f->f.ncomp == 3;
  if (hprec == QAMD_F64)
    hipLaunchKernelGGL((c3 ? k_import<double, 3> : k_import<double, 12>), grid, dim3(256), 0, stream(), f->f,
                       (const double *)stage, m);
  else
    hipLaunchKernelGGL((c3 ? k_import<float, 3> : k_import<float, 12>), grid, dim3(256), 0, stream(), f->f,
                       (const float *)stage, m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("import kernel: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

// prefill: the staging buffer first receives the host array, so that host
// elements the export does not write (the other parity of a lexicographic
// array) keep their values
static int export_common(const qamd_field *f, void *host, size_t bytes, int hprec, SrcMap m, int prefill) {
  void *stage = stage_dev(bytes);
  if (!stage || (prefill && h2d_pageable(stage, host, bytes))) return -1;
  dim3 grid = site_grid(f->lat->g, f->f.nrhs);
  const bool c3 = f->f.ncomp == 3;
  if (hprec == QAMD_F64)
    hipLaunchKernelGGL((c3 ? k_export<double, 3> : k_export<double, 12>), grid, dim3(256), 0, stream(), f->f,
                       (double *)stage, m);
  else
    hipLaunchKernelGGL((c3 ? k_export<float, 3> : k_export<float, 12>), grid, dim3(256), 0, stream(), f->f,
                       (float *)stage, m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("export kernel: %s", hipGetErrorString(e));
    return -1;
  }
  return d2h_pageable(host, stage, bytes);
}

int qamd_field_from_lex(qamd_field *f, const double *host) {
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 0;
  m.nreal = 2 * f->f.ncomp;
  size_t bytes = (size_t)f->f.nrhs * f->lat->g.V * m.nreal * sizeof(double);
  return import_common(f, host, bytes, QAMD_F64, m);
}

int qamd_field_to_lex(const qamd_field *f, double *host) {
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 0;
  m.nreal = 2 * f->f.ncomp;
  size_t bytes = (size_t)f->f.nrhs * f->lat->g.V * m.nreal * sizeof(double);
  // sites of the other parity keep the caller's values
  return export_common(f, host, bytes, QAMD_F64, m, 1);
}

static int check_qll(const qamd_field *f, const qamd_qll_layout *l) {
  for (int i = 0; i < 4; i++)
    if (l->physGeom[i] != f->lat->g.L[i] || l->innerGeom[i] * l->outerGeom[i] != l->physGeom[i]) {
      set_error("QLL layout does not match the lattice");
      return -1;
    }
  if (f->f.nrhs != 1) {
    set_error("QLL layout transfer needs nrhs = 1");
    return -1;
  }
  return 0;
}

int qamd_field_from_qll(qamd_field *f, const qamd_qll_layout *l, const void *host, int hprec) {
  if (check_qll(f, l)) return -1;
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 1;
  m.nreal = 2 * f->f.ncomp;
  m.lay = to_lay(l);
  size_t esz = hprec == QAMD_F64 ? 8 : 4;
  size_t half = (size_t)m.nreal * l->nSitesInner * l->nEvenOuter * esz;
  const char *h = (const char *)host + (f->f.parity ? half : 0);
  size_t bytes = (size_t)m.nreal * l->nSitesInner *
                 (f->f.parity ? (l->nSitesOuter - l->nEvenOuter) : l->nEvenOuter) * esz;
  return import_common(f, h, bytes, hprec, m);
}

int qamd_field_to_qll(const qamd_field *f, const qamd_qll_layout *l, void *host, int hprec) {
  if (check_qll(f, l)) return -1;
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 1;
  m.nreal = 2 * f->f.ncomp;
  m.lay = to_lay(l);
  size_t esz = hprec == QAMD_F64 ? 8 : 4;
  size_t half = (size_t)m.nreal * l->nSitesInner * l->nEvenOuter * esz;
  char *h = (char *)host + (f->f.parity ? half : 0);
  size_t bytes = (size_t)m.nreal * l->nSitesInner *
                 (f->f.parity ? (l->nSitesOuter - l->nEvenOuter) : l->nEvenOuter) * esz;
  return export_common(f, h, bytes, hprec, m, 0);
}

static int sites_xfer(qamd_field *f, const int *coords, int64_t nsites, void *host, int hprec, int to_host) {
  if (f->f.nrhs != 1) {
    set_error("site-table transfer needs nrhs = 1");
    return -1;
  }
  if (hprec != QAMD_F64 && hprec != QAMD_F32) {
    set_error("site-table transfer: host precision must be QAMD_F64 or QAMD_F32");
    return -1;
  }
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 2;
  m.nreal = 2 * f->f.ncomp;
  int *inv = nullptr;
  if (site_table_inverse(f->lat, coords, nsites, f->f.parity, &inv)) return -1;
  m.inv = inv;
  size_t bytes = (size_t)nsites * m.nreal * (hprec == QAMD_F64 ? 8 : 4);
  // export: sites of the other parity keep the caller's values
  int rc = to_host ? export_common(f, host, bytes, hprec, m, 1) : import_common(f, host, bytes, hprec, m);
  (void)hipStreamSynchronize(stream());   // the import kernel reads inv
  dev_free(inv);
  return rc;
}

int qamd_field_from_sites(qamd_field *f, const int *coords, int64_t nsites, const void *host, int hprec) {
  return sites_xfer(f, coords, nsites, (void *)host, hprec, 0);
}

int qamd_field_to_sites(const qamd_field *f, const int *coords, int64_t nsites, void *host, int hprec) {
  return sites_xfer((qamd_field *)f, coords, nsites, host, hprec, 1);
}

int qamd_field_gaussian(qamd_field *f, uint64_t seed, uint64_t strm) {
  hipLaunchKernelGGL((f->f.ncomp == 3 ? k_gaussian<3> : k_gaussian<12>), site_grid(f->lat->g, f->f.nrhs), dim3(256),
                     0, stream(), f->f, seed, strm);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int qamd_field_norm2(const qamd_field *f, double *out) {
  for (int r = 0; r < f->f.nrhs; r++)
    if (reduce_field(f->f, nullptr, r, &out[r])) return -1;
  return 0;
}

int qamd_field_redot(const qamd_field *a, const qamd_field *b, double *out) {
  if (a->f.ncomp != b->f.ncomp || a->f.nrhs != b->f.nrhs || a->lat->g.V != b->lat->g.V) {
    set_error("redot: shape mismatch");
    return -1;
  }
  for (int r = 0; r < a->f.nrhs; r++)
    if (reduce_field(a->f, &b->f, r, &out[r])) return -1;
  return 0;
}

int qamd_field_axpby(double a, const qamd_field *x, double b, qamd_field *y) {
  if (x->f.ncomp != y->f.ncomp || x->f.nrhs != y->f.nrhs || x->lat->g.V != y->lat->g.V) {
    set_error("axpby: shape mismatch");
    return -1;
  }
  hipLaunchKernelGGL((y->f.ncomp == 3 ? k_axpby<3> : k_axpby<12>), site_grid(y->lat->g, y->f.nrhs), dim3(256), 0,
                     stream(), a, x->f, b, y->f);
  QAMD_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
