// Dslash link sets: allocation, import from host (lexicographic or QLL
// layout), doublestore from device gauge fields, precision conversion.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_site.h"
#include "qamd_kernels.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

namespace qamd {

// storage slots per site: the term count, except forward-only staggered sets
// (kind 2), which always hold 8 slots (fat mu in slot mu, long mu in slot
// 4 + mu; without Naik terms the long slots stay zero)
static int nslots(const qamd_links *l) { return l->kind == 2 ? 8 : l->nterms; }
static int64_t link_elems(const qamd_links *l) { return l->lat->g.vh_pad * nslots(l) * 9; }

template <class S> __device__ inline S mk(double2 v);
template <> __device__ inline double2 mk<double2>(double2 v) { return v; }
template <> __device__ inline float2 mk<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

// The block layout is a permutation of the block's nterm x 9 elements, and
// with the rebuildable terms' third rows left out the rest is a prefix.
template <int G>
constexpr bool link_rows_ok(int nterm) {
  bool seen[16 * 9] = {};
  const int prefix = (nterm / G) * 9 - (8 / G) * 3;
  for (int sl = 0; sl < nterm; sl++)
    for (int k = 0; k < 9; k++) {
      const int r = link_row<G>(sl, k, nterm), i = r * G + sl % G;
      if (i < 0 || i >= nterm * 9 || seen[i]) return false;
      seen[i] = true;
      if ((sl >= nterm - 8 && k % 3 == 2) != (r >= prefix)) return false;
    }
  return true;
}
static_assert(link_rows_ok<1>(16) && link_rows_ok<2>(16) && link_rows_ok<4>(16) && link_rows_ok<1>(8) &&
                  link_rows_ok<2>(8) && link_rows_ok<4>(8),
              "link block rows");

// The forward-only layout is a permutation of the block's 8 x 9 elements; a
// slot's first two rows fill the leading positions of its words.
template <int G>
constexpr bool fb_rows_ok() {
  bool seen[8 * 9] = {};
  for (int sl = 0; sl < 8; sl++)
    for (int k = 0; k < 9; k++) {
      const int r = fb_row<G>(sl, k), q = fb_pos<G>(sl, k), i = r * G + q;
      if (q < 0 || q >= G || i < 0 || i >= 8 * 9 || seen[i]) return false;
      seen[i] = true;
      if (G > 1 && k != 8 && r / (G == 2 ? 4 : 2) != sl) return false;
      if (G == 2 && k % 3 != 2 && fb_row<G>(sl, k) % 4 == 3) return false;
      if (G == 4 && k % 3 != 2 && k >= 6 && q >= 2) return false;
    }
  return true;
}
static_assert(fb_rows_ok<1>() && fb_rows_ok<2>() && fb_rows_ok<4>(), "forward-only link rows");

// one term from a host-ordered staging array into the blocked layout
template <class H, class S>
__global__ void k_link_import(S *d0, S *d1, const H *src, Geom g, int t, int nterm, bool fb, SrcMap m) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int par = blockIdx.y;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  int64_t base, stride;
  if (m.kind == 0) {
    base = (int64_t)s.lex * 18;
    stride = 1;
  } else {
    int oi, ii;
    qll_index(m.lay, s.c, oi, ii);
    base = (int64_t)18 * m.lay.NI * oi + ii;
    stride = m.lay.NI;
  }
  S *d = par ? d1 : d0;
  for (int e = 0; e < 9; e++)
    d[link_idx<S>(cb, t, e, nterm, fb, g.vh_pad)] =
        mk<S>(make_double2((double)src[base + 2 * e * stride], (double)src[base + (2 * e + 1) * stride]));
}

// forward-only staggered set from FP64 gauge fields: slot mu = fat U_mu(x),
// slot 4 + mu = long L_mu(x) (zero without long links), at the sites of
// each parity
template <class S>
__global__ void k_fb_from_gauge(S *d0, S *d1, const double2 *fat, const double2 *lng, Geom g) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int par = blockIdx.y;
  if (cb >= g.vh) return;
  S *d = par ? d1 : d0;
  const int64_t vp = g.vh_pad;
  for (int mu = 0; mu < 4; mu++) {
    const int64_t o = ((int64_t)(mu * 2 + par) * 9) * vp + cb;
    for (int e = 0; e < 9; e++) {
      d[link_idx<S>(cb, mu, e, 8, true, vp)] = mk<S>(fat[o + e * vp]);
      d[link_idx<S>(cb, 4 + mu, e, 8, true, vp)] = mk<S>(lng ? lng[o + e * vp] : make_double2(0.0, 0.0));
    }
  }
}

// doublestore from FP64 gauge fields: term for (k, naik) at output parity par
template <class S>
__global__ void k_link_from_gauge(S *d0, S *d1, const double2 *fat, const double2 *lng, Geom g,
                                  int nterm, double sgn) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int par = blockIdx.y;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  S *d = par ? d1 : d0;
  const int64_t vp = g.vh_pad;
  for (int k = 0; k < 8; k++) {
    int mu = k >> 1, bwd = k & 1;
    for (int n = 0; n < (nterm == 16 ? 2 : 1); n++) {
      const double2 *src = n ? lng : fat;
      int t = (nterm == 16) ? 2 * k + n : k;
      if (!bwd) {
        const double2 *p = src + ((int64_t)(mu * 2 + par) * 9) * vp + cb;
        for (int e = 0; e < 9; e++) d[link_idx<S>(cb, t, e, nterm, false, vp)] = mk<S>(p[e * vp]);
      } else {
        int hop = n ? 3 : 1;
        int nb = nbr_cb(g, s, mu, -hop);
        const double2 *p = src + ((int64_t)(mu * 2 + (1 - par)) * 9) * vp + nb;
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            double2 v = p[(3 * r + c) * vp];   // element (c, r) of the source
            d[link_idx<S>(cb, t, 3 * c + r, nterm, false, vp)] = mk<S>(make_double2(sgn * v.x, -sgn * v.y));
          }
      }
    }
  }
}

template <class S> __device__ inline double2 ld2(const S *p, int64_t i);
template <> __device__ inline double2 ld2<double2>(const double2 *p, int64_t i) { return p[i]; }
template <> __device__ inline double2 ld2<float2>(const float2 *p, int64_t i) {
  return make_double2(p[i].x, p[i].y);
}
template <> __device__ inline double2 ld2<__half2>(const __half2 *p, int64_t i) {
  float2 f = __half22float2(p[i]);
  return make_double2(f.x, f.y);
}
template <> __device__ inline double2 ld2<short2>(const short2 *p, int64_t i) {   // stored / scale
  return make_double2(p[i].x / (double)Q16_ONE, p[i].y / (double)Q16_ONE);
}

// per-term max |component| as float bits (non-negative floats order as
// uints).  Each wavefront reduces one 64-site block (all terms) in registers
// and issues one atomic per term.
template <class S>
__global__ void k_link_absmax(const S *src, int64_t nblk, int nterm, bool fb, unsigned *mx) {
  int lane = threadIdx.x & 63;
  int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  float m[16];
  for (int t = 0; t < 16; t++) m[t] = 0.f;
  for (int64_t b = wave; b < nblk; b += nwaves)
    for (int t = 0; t < nterm; t++)
      for (int e = 0; e < 9; e++) {
        double2 v = ld2<S>(src, link_idx<S>(b * 64 + lane, t, e, nterm, fb, nblk * 64));
        m[t] = fmaxf(m[t], (float)fmax(fabs(v.x), fabs(v.y)));
      }
  for (int t = 0; t < nterm; t++) {
    float v = m[t];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if (lane == 0) atomicMax(&mx[t], __float_as_uint(v));
  }
}

template <class SI, class SO>
__global__ void k_link_convert(SO *dst, const SI *src, int64_t n, int nterm, bool fb, const float *iscale,
                               const float *oscale) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    // (site block, term, element, lane) of work item i; consecutive threads
    // take consecutive sites
    int64_t q = i >> 6;
    const int lane = (int)(i & 63), k = (int)(q % 9);
    q /= 9;
    const int t = (int)(q % nterm);
    const int64_t cb = (q / nterm) * 64 + lane;
    const int64_t vp = n / ((int64_t)nterm * 9);   // n covers every element of one parity's set
    const int64_t o = link_idx<SO>(cb, t, k, nterm, fb, vp);
    double2 v = ld2<SI>(src, link_idx<SI>(cb, t, k, nterm, fb, vp));
    double f = (double)iscale[t] / (double)oscale[t];
    v.x *= f;
    v.y *= f;
    if constexpr (std::is_same<SO, __half2>::value) dst[o] = __floats2half2_rn((float)v.x, (float)v.y);
    else if constexpr (std::is_same<SO, short2>::value) dst[o] = q16_pack((float)v.x, (float)v.y);
    else dst[o] = mk<SO>(v);
  }
}

// Terms of the form c * W with W unitary (c > 0, det W = +-1): then the
// third row is det W * conj(row0 x row1) / c.  Checks the nk terms
// t0 + dt*k of FP64 links (16-term staggered: the eight long terms 2k+1;
// Wilson: all eight; forward-only staggered: the four long slots 4 + mu) to
// 1e-12 relative, records det W per (site, k) and flags any term that is not
// of that form.
__global__ void k_long_check(const double2 *d, uint8_t *sign, Geom g, int nterm, bool fb, int t0, int dt, int nk,
                             unsigned *bad) {
  const int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= g.vh) return;
  unsigned bits = 0, fail = 0;
  for (int k = 0; k < nk; k++) {
    double2 m[9];
    for (int e = 0; e < 9; e++) m[e] = d[link_idx<double2>(cb, t0 + dt * k, e, nterm, fb, g.vh_pad)];
    // row r holds elements 3c + r
    double c2 = 0, n2 = 0;
    double2 dot = make_double2(0, 0);
    for (int c = 0; c < 3; c++) {
      const double2 a = m[3 * c], b = m[3 * c + 1];
      c2 += a.x * a.x + a.y * a.y;
      n2 += b.x * b.x + b.y * b.y;
      dot.x += a.x * b.x + a.y * b.y;   // row0 . conj(row1)
      dot.y += a.y * b.x - a.x * b.y;
    }
    const double tol = 1e-12;
    if (!(c2 > 1e-200) || fabs(n2 - c2) > tol * c2 || sqrt(dot.x * dot.x + dot.y * dot.y) > tol * c2) {
      fail = 1;
      continue;
    }
    const double ic = 1.0 / sqrt(c2);
    double2 x[3];   // conj(row0 x row1) / c
    double proj = 0, dev = 0;
    for (int c = 0; c < 3; c++) {
      const double2 a1 = m[3 * ((c + 1) % 3)], a2 = m[3 * ((c + 2) % 3)];
      const double2 b1 = m[3 * ((c + 1) % 3) + 1], b2 = m[3 * ((c + 2) % 3) + 1];
      const double re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
      const double im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
      x[c] = make_double2(re * ic, -im * ic);
      proj += m[3 * c + 2].x * x[c].x + m[3 * c + 2].y * x[c].y;
    }
    const double sg = proj < 0 ? -1.0 : 1.0;
    for (int c = 0; c < 3; c++) {
      const double dx = m[3 * c + 2].x - sg * x[c].x, dy = m[3 * c + 2].y - sg * x[c].y;
      dev += dx * dx + dy * dy;
    }
    if (dev > tol * tol * c2) fail = 1;
    if (sg < 0) bits |= 1u << k;
  }
  sign[cb] = (uint8_t)bits;
  if (fail) atomicOr(bad, 1u);
}

// link sets whose Dslash may rebuild third rows: 16-term staggered (the long
// terms), Wilson (every term) and forward-only staggered with long links
static bool rebuild_kind(const qamd_links *l) {
  return (l->kind == 0 && l->nterms == 16) || l->kind == 1 || (l->kind == 2 && l->nterms == 8);
}

// after FP64 links change: decide whether the Dslash may rebuild the third
// rows (every checked term passes)
static uint64_t g_links_ver = 0;
static void new_version(qamd_links *l) { l->ver = __atomic_add_fetch(&g_links_ver, 1, __ATOMIC_RELAXED); }

static int long_link_check(qamd_links *l) {
  new_version(l);
  l->recon = 0;
  if (!rebuild_kind(l) || l->prec != QAMD_F64) return 0;
  const int t0 = l->kind == 0 ? 1 : l->kind == 2 ? 4 : 0, dt = l->kind == 0 ? 2 : 1, nk = l->kind == 2 ? 4 : 8;
  const Geom &g = l->lat->g;
  for (int p = 0; p < 2; p++)
    if (!l->lsign[p] && !(l->lsign[p] = (uint8_t *)dev_alloc(g.vh_pad))) return -1;
  unsigned *bad = (unsigned *)dev_alloc(sizeof(unsigned));
  if (!bad) return -1;
  unsigned hb = 0;
  hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned), stream());
  for (int p = 0; p < 2 && e == hipSuccess; p++) {
    hipLaunchKernelGGL(k_long_check, dim3((unsigned)((g.vh + 255) / 256)), dim3(256), 0, stream(),
                       (const double2 *)l->data[p], l->lsign[p], g, nslots(l), l->kind == 2, t0, dt, nk, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hb, bad, sizeof hb, hipMemcpyDeviceToHost, stream());
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  dev_free(bad);
  if (e != hipSuccess) {
    set_error("long-link check: %s", hipGetErrorString(e));
    return -1;
  }
  l->recon = hb ? 0 : 1;
  return 0;
}

// A lower-precision copy of links that pass the check rebuilds its long
// terms' third rows the same way (same determinant signs): its stored rows
// are the roundings of a scaled unitary matrix's, so the rebuilt row carries
// the storage precision's error, as a stored one would.
static int inherit_recon(qamd_links *dst, const qamd_links *src) {
  new_version(dst);
  dst->recon = 0;
  if (!src->recon || !rebuild_kind(dst)) return 0;
  const Geom &g = dst->lat->g;
  for (int p = 0; p < 2; p++) {
    if (!dst->lsign[p] && !(dst->lsign[p] = (uint8_t *)dev_alloc(g.vh_pad))) return -1;
    if (hipMemcpyAsync(dst->lsign[p], src->lsign[p], g.vh_pad, hipMemcpyDeviceToDevice, stream()) != hipSuccess) {
      set_error("links_convert: sign copy failed");
      return -1;
    }
  }
  dst->recon = 1;
  return 0;
}

}  // namespace qamd

using namespace qamd;

extern "C" {

qamd_links *qamd_links_create(qamd_lattice *lat, int kind, int prec, int nterms) {
  if (!lat) { set_error("links_create: NULL lattice"); return nullptr; }
  if (kind == 0 && nterms != 8 && nterms != 16) { set_error("staggered links need 8 or 16 terms"); return nullptr; }
  if (kind == 1 && nterms != 8) { set_error("Wilson links need 8 terms"); return nullptr; }
  if (kind == 2 && nterms != 4 && nterms != 8) {
    set_error("forward-only staggered links need 4 (fat) or 8 (fat + long) terms");
    return nullptr;
  }
  if (kind < 0 || kind > 2) { set_error("links_create: bad kind"); return nullptr; }
  if (prec < QAMD_F64 || prec > QAMD_Q16) { set_error("links_create: bad precision"); return nullptr; }
  qamd_links *l = (qamd_links *)calloc(1, sizeof(qamd_links));
  l->lat = lat;
  l->kind = kind;
  l->prec = prec;
  l->nterms = nterms;
  for (int t = 0; t < 16; t++) l->scale[t] = 1.0f;
  int64_t n = link_elems(l);
  for (int p = 0; p < 2; p++) {
    l->data[p] = dev_alloc(n * prec_bytes(prec));
    if (!l->data[p]) {
      qamd_links_destroy(l);
      return nullptr;
    }
    if (hipMemsetAsync(l->data[p], 0, n * prec_bytes(prec), stream()) != hipSuccess) {
      set_error("links_create: memset failed");
      qamd_links_destroy(l);
      return nullptr;
    }
  }
  return l;
}

void qamd_links_destroy(qamd_links *l) {
  if (!l) return;
  (void)hipStreamSynchronize(stream());
  dev_free(l->data[0]);
  dev_free(l->data[1]);
  dev_free(l->lsign[0]);
  dev_free(l->lsign[1]);
  free(l);
}

// import with an FP64/FP32 destination; 16-bit formats go through FP64
static int import_terms(qamd_links *l, const void *const terms[], int hprec, SrcMap m, size_t term_bytes) {
  qamd_links *dst = l;
  if (l->prec == QAMD_F16 || l->prec == QAMD_Q16) {
    dst = qamd_links_create(l->lat, l->kind, QAMD_F64, l->nterms);
    if (!dst) return -1;
  }
  void *stage = stage_dev(term_bytes);
  if (!stage) {
    if (dst != l) qamd_links_destroy(dst);
    return -1;
  }
  const Geom &g = l->lat->g;
  dim3 grid((unsigned)((g.vh + 255) / 256), 2);
  int rc = 0;
  for (int t = 0; t < l->nterms && !rc; t++) {
    // stream order: the upload of term t + 1 waits for the import kernel of t
    if (h2d_pageable(stage, terms[t], term_bytes)) {
      rc = -1;
      break;
    }
    const int ns = nslots(l);
    const bool fb = l->kind == 2;
    if (dst->prec == QAMD_F64) {
      if (hprec == QAMD_F64)
        hipLaunchKernelGGL((k_link_import<double, double2>), grid, dim3(256), 0, stream(), (double2 *)dst->data[0],
                           (double2 *)dst->data[1], (const double *)stage, g, t, ns, fb, m);
      else
        hipLaunchKernelGGL((k_link_import<float, double2>), grid, dim3(256), 0, stream(), (double2 *)dst->data[0],
                           (double2 *)dst->data[1], (const float *)stage, g, t, ns, fb, m);
    } else {
      if (hprec == QAMD_F64)
        hipLaunchKernelGGL((k_link_import<double, float2>), grid, dim3(256), 0, stream(), (float2 *)dst->data[0],
                           (float2 *)dst->data[1], (const double *)stage, g, t, ns, fb, m);
      else
        hipLaunchKernelGGL((k_link_import<float, float2>), grid, dim3(256), 0, stream(), (float2 *)dst->data[0],
                           (float2 *)dst->data[1], (const float *)stage, g, t, ns, fb, m);
    }
    if (hipGetLastError() != hipSuccess) {
      set_error("links import kernel failed");
      rc = -1;
    }
  }
  if (!rc && hipStreamSynchronize(stream()) != hipSuccess) {
    set_error("links import kernel failed");
    rc = -1;
  }
  if (dst != l) {   // 16-bit: checked in FP64, then converted (the copy inherits the result)
    if (!rc) rc = long_link_check(dst);
    if (!rc) rc = qamd_links_convert(l, dst);
    qamd_links_destroy(dst);
  } else if (!rc) {
    rc = long_link_check(l);
  }
  return rc;
}

int qamd_links_from_lex(qamd_links *l, const double *host) {
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 0;
  const void *terms[16];
  size_t tb = (size_t)l->lat->g.V * 18 * sizeof(double);
  for (int t = 0; t < l->nterms; t++) terms[t] = (const char *)host + t * tb;
  return import_terms(l, terms, QAMD_F64, m, tb);
}

int qamd_links_from_qll(qamd_links *l, const qamd_qll_layout *lay, const void *const u[],
                        const void *const u3[], int hprec) {
  for (int i = 0; i < 4; i++)
    if (lay->physGeom[i] != l->lat->g.L[i]) {
      set_error("links_from_qll: layout does not match lattice");
      return -1;
    }
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 1;
  m.lay = to_lay(lay);
  const void *terms[16];
  if (l->kind == 2) {   // forward links only: QLL's u[2mu] (and u3[2mu])
    if (l->nterms == 8 && !u3) {
      set_error("links_from_qll: forward-only links with Naik terms need u3");
      return -1;
    }
    for (int mu = 0; mu < 4; mu++) {
      terms[mu] = u[2 * mu];
      if (l->nterms == 8) terms[4 + mu] = u3[2 * mu];
    }
  } else if (l->nterms == 16) {
    if (!u3) {
      set_error("links_from_qll: 16 terms need u3");
      return -1;
    }
    for (int k = 0; k < 8; k++) {
      terms[2 * k] = u[k];
      terms[2 * k + 1] = u3[k];
    }
  } else {
    for (int k = 0; k < 8; k++) terms[k] = u[k];
  }
  size_t tb = (size_t)l->lat->g.V * 18 * (hprec == QAMD_F64 ? 8 : 4);
  return import_terms(l, terms, hprec, m, tb);
}

int qamd_links_from_gauge(qamd_links *l, const qamd_gauge *fat, const qamd_gauge *lng) {
  if ((l->nterms == 16 || (l->kind == 2 && l->nterms == 8)) && !lng) {
    set_error("links_from_gauge: 16 terms need long links");
    return -1;
  }
  qamd_links *dst = l;
  if (l->prec == QAMD_F16 || l->prec == QAMD_Q16) {
    dst = qamd_links_create(l->lat, l->kind, QAMD_F64, l->nterms);
    if (!dst) return -1;
  }
  const Geom &g = l->lat->g;
  dim3 grid((unsigned)((g.vh + 255) / 256), 2);
  double sgn = l->kind == 0 ? -1.0 : 1.0;
  const double2 *lp = lng ? lng->d : fat->d;
  if (l->kind == 2) {
    const double2 *lf = l->nterms == 8 ? lng->d : nullptr;
    if (dst->prec == QAMD_F64)
      hipLaunchKernelGGL(k_fb_from_gauge<double2>, grid, dim3(256), 0, stream(), (double2 *)dst->data[0],
                         (double2 *)dst->data[1], fat->d, lf, g);
    else
      hipLaunchKernelGGL(k_fb_from_gauge<float2>, grid, dim3(256), 0, stream(), (float2 *)dst->data[0],
                         (float2 *)dst->data[1], fat->d, lf, g);
  } else if (dst->prec == QAMD_F64)
    hipLaunchKernelGGL(k_link_from_gauge<double2>, grid, dim3(256), 0, stream(), (double2 *)dst->data[0],
                       (double2 *)dst->data[1], fat->d, lp, g, l->nterms, sgn);
  else
    hipLaunchKernelGGL(k_link_from_gauge<float2>, grid, dim3(256), 0, stream(), (float2 *)dst->data[0],
                       (float2 *)dst->data[1], fat->d, lp, g, l->nterms, sgn);
  int rc = 0;
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream()) != hipSuccess) {
    set_error("links_from_gauge kernel failed");
    rc = -1;
  }
  if (dst != l) {   // 16-bit: checked in FP64, then converted (the copy inherits the result)
    if (!rc) rc = long_link_check(dst);
    if (!rc) rc = qamd_links_convert(l, dst);
    qamd_links_destroy(dst);
  } else if (!rc) {
    rc = long_link_check(l);
  }
  return rc;
}

int qamd_links_reconstructs(const qamd_links *l) { return l && l->recon ? 1 : 0; }

size_t qamd_links_bytes(const qamd_links *l) {
  if (!l) return 0;
  size_t n = 2 * (size_t)link_elems(l) * prec_bytes(l->prec);
  for (int p = 0; p < 2; p++)
    if (l->lsign[p]) n += (size_t)l->lat->g.vh_pad;
  return n;
}

int qamd_links_convert(qamd_links *dst, const qamd_links *src) {
  if (dst->lat->g.V != src->lat->g.V || dst->nterms != src->nterms || dst->kind != src->kind) {
    set_error("links_convert: mismatch");
    return -1;
  }
  int64_t n = link_elems(src);
  int64_t nblk = src->lat->g.vh_pad / 64;
  float *sc = (float *)dev_alloc(64 * sizeof(float));
  unsigned *mx = (unsigned *)dev_alloc(16 * sizeof(unsigned));
  if (!sc || !mx) {
    dev_free(sc);
    dev_free(mx);
    return -1;
  }
  float oscale[16];
  for (int t = 0; t < 16; t++) oscale[t] = 1.0f;
  if (dst->prec == QAMD_F16 || dst->prec == QAMD_Q16) {
    QAMD_HIP(hipMemsetAsync(mx, 0, 16 * sizeof(unsigned), stream()));
    for (int p = 0; p < 2; p++) {
      if (src->prec == QAMD_F64)
        hipLaunchKernelGGL(k_link_absmax<double2>, dim3(2048), dim3(256), 0, stream(), (const double2 *)src->data[p], nblk, nslots(src), src->kind == 2, mx);
      else if (src->prec == QAMD_F32)
        hipLaunchKernelGGL(k_link_absmax<float2>, dim3(2048), dim3(256), 0, stream(), (const float2 *)src->data[p], nblk, nslots(src), src->kind == 2, mx);
      else if (src->prec == QAMD_F16)
        hipLaunchKernelGGL(k_link_absmax<__half2>, dim3(2048), dim3(256), 0, stream(), (const __half2 *)src->data[p], nblk, nslots(src), src->kind == 2, mx);
      else
        hipLaunchKernelGGL(k_link_absmax<short2>, dim3(2048), dim3(256), 0, stream(), (const short2 *)src->data[p], nblk, nslots(src), src->kind == 2, mx);
    }
    unsigned hm[16];
    QAMD_HIP(hipMemcpyAsync(hm, mx, sizeof hm, hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    for (int t = 0; t < nslots(src); t++) {
      float m;
      memcpy(&m, &hm[t], 4);
      m *= src->scale[t];
      oscale[t] = m > 0 ? m : 1.0f;
    }
  }
  QAMD_HIP(hipMemcpyAsync(sc, src->scale, 16 * sizeof(float), hipMemcpyHostToDevice, stream()));
  QAMD_HIP(hipMemcpyAsync(sc + 16, oscale, 16 * sizeof(float), hipMemcpyHostToDevice, stream()));
  for (int p = 0; p < 2; p++) {
#define CONV(SI, SO)                                                                                   \
  hipLaunchKernelGGL((k_link_convert<SI, SO>), dim3(2048), dim3(256), 0, stream(), (SO *)dst->data[p], \
                     (const SI *)src->data[p], n, nslots(src), src->kind == 2, sc, sc + 16)
    int a = src->prec, b = dst->prec;
    if (a == QAMD_F64 && b == QAMD_F64) CONV(double2, double2);
    else if (a == QAMD_F64 && b == QAMD_F32) CONV(double2, float2);
    else if (a == QAMD_F64 && b == QAMD_F16) CONV(double2, __half2);
    else if (a == QAMD_F32 && b == QAMD_F64) CONV(float2, double2);
    else if (a == QAMD_F32 && b == QAMD_F32) CONV(float2, float2);
    else if (a == QAMD_F32 && b == QAMD_F16) CONV(float2, __half2);
    else if (a == QAMD_F16 && b == QAMD_F64) CONV(__half2, double2);
    else if (a == QAMD_F16 && b == QAMD_F32) CONV(__half2, float2);
    else if (a == QAMD_F16 && b == QAMD_F16) CONV(__half2, __half2);
    else if (a == QAMD_F64 && b == QAMD_Q16) CONV(double2, short2);
    else if (a == QAMD_F32 && b == QAMD_Q16) CONV(float2, short2);
    else if (a == QAMD_Q16 && b == QAMD_F64) CONV(short2, double2);
    else if (a == QAMD_Q16 && b == QAMD_F32) CONV(short2, float2);
    else if (a == QAMD_Q16 && b == QAMD_Q16) CONV(short2, short2);
    else {
      set_error("links_convert: unsupported precision pair %d -> %d", a, b);
      dev_free(sc);
      dev_free(mx);
      return -1;
    }
#undef CONV
  }
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipStreamSynchronize(stream()));
  for (int t = 0; t < 16; t++) dst->scale[t] = oscale[t];
  dev_free(sc);
  dev_free(mx);
  return dst->prec == QAMD_F64 ? long_link_check(dst) : inherit_recon(dst, src);
}

}  // extern "C"
