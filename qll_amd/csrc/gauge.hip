// Gauge fields on the device: seeded SU(3) generation, staggered phases,
// fat7 + Naik smearing as a staple-product kernel, plaquette, host transfer.
//
// fat7 (lib/gauge.c:188-416) is evaluated per (site, direction) as the sum of
// nested staples, all in FP64:
//   St_nu[X](x) = U_nu(x) X(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ X(x-nu) U_nu(x-nu+mu)
//   fat_mu = (c1 - 6 cL) U_mu + c3 sum_nu St_nu[U_mu] + cL sum_nu St_nu[St_nu[U_mu]]
//          + c5 sum_{nu,rho} St_rho[St_nu[U_mu]]
//          + c7 sum_{nu,rho,sig} St_sig[St_rho[St_nu[U_mu]]]
// which is the same sum of path products QLL's genStaple recursion builds with
// transporters and temporary fields; here every site recomputes its own
// paths from U (no temporaries, one pass, gathers served by L2/MALL).
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_site.h"
#include "qamd_kernels.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

namespace qamd {

struct M3 {
  double2 e[9];   // column major: (r,c) at 3c+r
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {   // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 ccmul(double2 a, double2 b) {   // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// C = A B, A^+ B, A B^+
__device__ inline M3 mul_nn(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 s = cmul(A.e[r], B.e[3 * c]);
      s = make_double2(s.x, s.y);
#pragma unroll
      for (int k = 1; k < 3; k++) {
        double2 t = cmul(A.e[3 * k + r], B.e[3 * c + k]);
        s.x += t.x;
        s.y += t.y;
      }
      C.e[3 * c + r] = s;
    }
  return C;
}
__device__ inline M3 mul_an(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double2 t = ccmul(A.e[3 * r + k], B.e[3 * c + k]);   // conj(A(k,r)) B(k,c)
        s.x += t.x;
        s.y += t.y;
      }
      C.e[3 * c + r] = s;
    }
  return C;
}
__device__ inline M3 mul_na(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double2 t = cmulc(A.e[3 * k + r], B.e[3 * k + c]);   // A(r,k) conj(B(c,k))
        s.x += t.x;
        s.y += t.y;
      }
      C.e[3 * c + r] = s;
    }
  return C;
}
__device__ inline void madd(M3 &A, double a, const M3 &B) {
#pragma unroll
  for (int i = 0; i < 9; i++) {
    A.e[i].x += a * B.e[i].x;
    A.e[i].y += a * B.e[i].y;
  }
}

struct GView {
  const double2 *d;
  Geom g;
  __device__ M3 load(int mu, const int c[4]) const {
    int lex = c[0] + g.stride[1] * c[1] + g.stride[2] * c[2] + g.stride[3] * c[3];
    int par = (c[0] + c[1] + c[2] + c[3]) & 1;
    const double2 *p = d + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + (lex >> 1);
    M3 m;
#pragma unroll
    for (int e = 0; e < 9; e++) m.e[e] = p[e * g.vh_pad];
    return m;
  }
  __device__ void shift(int c[4], int mu, int h) const {
    int v = c[mu] + h;
    v += (v < 0) ? g.L[mu] : 0;
    v -= (v >= g.L[mu]) ? g.L[mu] : 0;
    c[mu] = v;
  }
};

// explicit depth unrolling (device recursion is avoided)
template <int LEVEL>
__device__ M3 nestedT(const GView &U, int mu, const int *dirs, const int c[4]);
template <>
__device__ M3 nestedT<0>(const GView &U, int mu, const int *, const int c[4]) { return U.load(mu, c); }
template <int LEVEL>
__device__ M3 nestedT(const GView &U, int mu, const int *dirs, const int c[4]) {
  int nu = dirs[LEVEL - 1];
  int b[4] = {c[0], c[1], c[2], c[3]};
  M3 Un = U.load(nu, b);
  U.shift(b, nu, 1);
  M3 t = mul_nn(Un, nestedT<LEVEL - 1>(U, mu, dirs, b));
  int d[4] = {c[0], c[1], c[2], c[3]};
  U.shift(d, mu, 1);
  M3 res = mul_na(t, U.load(nu, d));
  int e[4] = {c[0], c[1], c[2], c[3]};
  U.shift(e, nu, -1);
  M3 t2 = mul_an(U.load(nu, e), nestedT<LEVEL - 1>(U, mu, dirs, e));
  U.shift(e, mu, 1);
  madd(res, 1.0, mul_nn(t2, U.load(nu, e)));
  return res;
}

struct Fat7Coef {
  double c1, c3, c5, c7, cL, naik;
};

// one thread per (site, mu): grid.x over checkerboard sites, grid.y = 2*mu+par
__global__ void __launch_bounds__(128) k_fat7(double2 *fl, const double2 *u, Geom g, Fat7Coef C) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  GView U{u, g};
  M3 F = U.load(mu, s.c);
#pragma unroll
  for (int i = 0; i < 9; i++) {
    F.e[i].x *= C.c1 - 6 * C.cL;
    F.e[i].y *= C.c1 - 6 * C.cL;
  }
  bool have5 = C.c5 != 0 || C.c7 != 0 || C.cL != 0;
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    int dirs[3] = {nu, 0, 0};
    if (C.c3 != 0 || have5) madd(F, C.c3, nestedT<1>(U, mu, dirs, s.c));
    if (C.cL != 0) {
      dirs[1] = nu;
      madd(F, C.cL, nestedT<2>(U, mu, dirs, s.c));
    }
    if (C.c5 != 0 || C.c7 != 0) {
      for (int rho = 0; rho < 4; rho++) {
        if (rho == mu || rho == nu) continue;
        dirs[1] = rho;
        madd(F, C.c5, nestedT<2>(U, mu, dirs, s.c));
        if (C.c7 != 0) {
          for (int sig = 0; sig < 4; sig++) {
            if (sig == mu || sig == nu || sig == rho) continue;
            dirs[2] = sig;
            madd(F, C.c7, nestedT<3>(U, mu, dirs, s.c));
          }
        }
      }
    }
  }
  double2 *p = fl + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) p[e * g.vh_pad] = F.e[e];
}

// long links: naik * U(x) U(x+mu) U(x+2mu)   (lib/gauge.c:369-392)
__global__ void k_longlink(double2 *ll, const double2 *u, Geom g, double naik) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  GView U{u, g};
  int c[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
  M3 A = U.load(mu, c);
  U.shift(c, mu, 1);
  A = mul_nn(A, U.load(mu, c));
  U.shift(c, mu, 1);
  A = mul_nn(A, U.load(mu, c));
  double2 *p = ll + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) p[e * g.vh_pad] = make_double2(naik * A.e[e].x, naik * A.e[e].y);
}

// Gram-Schmidt in the column order of makeUsite (lib/vla.c:1635-1654), then
// divide by the principal cube root of the determinant
__device__ void su3_from_gauss(M3 &m) {
  for (int i = 0; i < 3; i++) {
    double n = 0;
    for (int r = 0; r < 3; r++) n += m.e[3 * i + r].x * m.e[3 * i + r].x + m.e[3 * i + r].y * m.e[3 * i + r].y;
    double ni = 1.0 / sqrt(n);
    for (int r = 0; r < 3; r++) { m.e[3 * i + r].x *= ni; m.e[3 * i + r].y *= ni; }
    // projections removed twice: one pass leaves eps * cond errors (~1e-12
    // for the worst of millions of Gaussian matrices), two leave a few eps
    for (int pass = 0; pass < 2; pass++)
      for (int j = 0; j < i; j++) {
        double2 d = make_double2(0, 0);
        for (int r = 0; r < 3; r++) {
          double2 t = ccmul(m.e[3 * j + r], m.e[3 * i + r]);
          d.x += t.x;
          d.y += t.y;
        }
        for (int r = 0; r < 3; r++) {
          double2 t = cmul(d, m.e[3 * j + r]);
          m.e[3 * i + r].x -= t.x;
          m.e[3 * i + r].y -= t.y;
        }
        n = 0;
        for (int r = 0; r < 3; r++) n += m.e[3 * i + r].x * m.e[3 * i + r].x + m.e[3 * i + r].y * m.e[3 * i + r].y;
        ni = 1.0 / sqrt(n);
        for (int r = 0; r < 3; r++) { m.e[3 * i + r].x *= ni; m.e[3 * i + r].y *= ni; }
      }
  }
  // det by cofactors of the first row: (r,c) = e[3c+r]
  double2 a = m.e[0], b = m.e[3], c = m.e[6], d = m.e[1], e = m.e[4], f = m.e[7], gg = m.e[2], h = m.e[5], k = m.e[8];
  double2 ek = cmul(e, k), fh = cmul(f, h), dk = cmul(d, k), fg = cmul(f, gg), dh = cmul(d, h), eg = cmul(e, gg);
  double2 t1 = cmul(a, make_double2(ek.x - fh.x, ek.y - fh.y));
  double2 t2 = cmul(b, make_double2(dk.x - fg.x, dk.y - fg.y));
  double2 t3 = cmul(c, make_double2(dh.x - eg.x, dh.y - eg.y));
  double2 det = make_double2(t1.x - t2.x + t3.x, t1.y - t2.y + t3.y);
  double ph = atan2(det.y, det.x);
  double sn, cs;
  sincos(-ph / 3, &sn, &cs);
  double2 fct = make_double2(cs, sn);
  for (int i = 0; i < 9; i++) m.e[i] = cmul(m.e[i], fct);
}

__global__ void k_random_su3(double2 *u, Geom g, uint64_t seed) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  M3 m;
  for (int k = 0; k < 9; k++) {
    double a, b;
    gauss2(seed, 16 + mu, (uint64_t)s.lex, k, a, b);
    m.e[k] = make_double2(a * 0.70710678118654752440, b * 0.70710678118654752440);
  }
  su3_from_gauss(m);
  double2 *p = u + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + cb;
  for (int e = 0; e < 9; e++) p[e * g.vh_pad] = m.e[e];
}

__global__ void k_stag_phases(double2 *u, Geom g) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  int acc = 0;
  for (int nu = 0; nu < mu; nu++) acc += s.c[nu];
  if (!(acc & 1)) return;
  double2 *p = u + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + cb;
  for (int e = 0; e < 9; e++) {
    double2 v = p[e * g.vh_pad];
    p[e * g.vh_pad] = make_double2(-v.x, -v.y);
  }
}

struct FinStore6 {
  double *dst;
  double scale;
  __device__ void operator()(const double *t) const {
    for (int i = 0; i < 6; i++) dst[i] = t[i] * scale;
  }
};

__global__ void k_plaq(const double2 *u, Geom g, RedScratch rs) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  GView U{u, g};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * g.vh; i += (int64_t)gridDim.x * blockDim.x) {
    int par = (int)(i / g.vh), cb = (int)(i % g.vh);
    Site s = site_of(g, cb, par);
    int l = 0;
    for (int mu = 0; mu < 4; mu++)
      for (int nu = 0; nu < mu; nu++) {
        int a[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
        int b[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
        U.shift(a, mu, 1);
        U.shift(b, nu, 1);
        M3 A = mul_nn(U.load(mu, s.c), U.load(nu, a));
        M3 B = mul_nn(U.load(nu, s.c), U.load(mu, b));
        M3 P = mul_na(A, B);
        acc[l++] += P.e[0].x + P.e[4].x + P.e[8].x;
      }
  }
  grid_partial<6>(acc, rs);
}

// host transfer: [4][V][18] lexicographic, 4 QLL arrays, or 4 arrays in a
// caller's site order; row_major: host matrices are stored row by row
// (device element e = 3c + r holds entry (r, c), column-major like QLL)
template <class H>
__global__ void k_gauge_xfer(double2 *u, H *h, Geom g, int mu, SrcMap m, int to_host, int row_major) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int par = blockIdx.y;
  if (cb >= g.vh) return;
  Site s = site_of(g, cb, par);
  int64_t base, stride;
  if (m.kind == 0) {
    base = ((int64_t)mu * g.V + s.lex) * 18;
    stride = 1;
  } else if (m.kind == 2) {
    base = (int64_t)m.inv[s.lex] * 18;
    stride = 1;
  } else {
    int oi, ii;
    qll_index(m.lay, s.c, oi, ii);
    base = (int64_t)18 * m.lay.NI * oi + ii;
    stride = m.lay.NI;
  }
  double2 *p = u + ((int64_t)(mu * 2 + par) * 9) * g.vh_pad + cb;
  for (int e = 0; e < 9; e++) {
    const int he = row_major ? 3 * (e % 3) + e / 3 : e;
    if (to_host) {
      double2 v = p[e * g.vh_pad];
      h[base + 2 * he * stride] = (H)v.x;
      h[base + (2 * he + 1) * stride] = (H)v.y;
    } else {
      p[e * g.vh_pad] = make_double2((double)h[base + 2 * he * stride], (double)h[base + (2 * he + 1) * stride]);
    }
  }
}

static dim3 gauge_grid(const Geom &g, int bs) { return dim3((unsigned)((g.vh + bs - 1) / bs), 8); }

}  // namespace qamd

using namespace qamd;

extern "C" {

qamd_gauge *qamd_gauge_create(qamd_lattice *lat) {
  if (!lat) { set_error("gauge_create: NULL lattice"); return nullptr; }
  qamd_gauge *u = (qamd_gauge *)calloc(1, sizeof(qamd_gauge));
  u->lat = lat;
  size_t n = (size_t)8 * 9 * lat->g.vh_pad;
  u->d = (double2 *)dev_alloc(n * sizeof(double2));
  if (!u->d || hipMemsetAsync(u->d, 0, n * sizeof(double2), stream()) != hipSuccess) {
    dev_free(u->d);
    free(u);
    return nullptr;
  }
  return u;
}

void qamd_gauge_destroy(qamd_gauge *u) {
  if (!u) return;
  (void)hipStreamSynchronize(stream());
  dev_free(u->d);
  free(u);
}

static int gauge_xfer_lex(qamd_gauge *u, double *host, int to_host) {
  const Geom &g = u->lat->g;
  size_t bytes = (size_t)4 * g.V * 18 * sizeof(double);
  double *st = (double *)stage_dev(bytes);
  if (!st) return -1;
  SrcMap m;
  memset(&m, 0, sizeof m);
  if (!to_host && h2d_pageable(st, host, bytes)) return -1;
  for (int mu = 0; mu < 4; mu++)
    hipLaunchKernelGGL(k_gauge_xfer<double>, dim3((unsigned)((g.vh + 255) / 256), 2), dim3(256), 0, stream(),
                       u->d, st, g, mu, m, to_host, 0);
  QAMD_HIP(hipGetLastError());
  if (to_host) return d2h_pageable(host, st, bytes);
  return 0;
}

int qamd_gauge_from_lex(qamd_gauge *u, const double *host) { return gauge_xfer_lex(u, (double *)host, 0); }
int qamd_gauge_to_lex(const qamd_gauge *u, double *host) { return gauge_xfer_lex((qamd_gauge *)u, host, 1); }

static int gauge_xfer_qll(qamd_gauge *u, const qamd_qll_layout *l, void *const host[4], int hprec, int to_host) {
  const Geom &g = u->lat->g;
  for (int i = 0; i < 4; i++)
    if (l->physGeom[i] != g.L[i]) {
      set_error("gauge QLL transfer: layout does not match lattice");
      return -1;
    }
  size_t esz = hprec == QAMD_F64 ? 8 : 4;
  size_t bytes = (size_t)g.V * 18 * esz;
  void *st = stage_dev(bytes);
  if (!st) return -1;
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 1;
  m.lay = to_lay(l);
  for (int mu = 0; mu < 4; mu++) {
    // stream order: the upload of direction mu + 1 waits for kernel mu
    if (!to_host && h2d_pageable(st, host[mu], bytes)) return -1;
    dim3 grid((unsigned)((g.vh + 255) / 256), 2);
    if (hprec == QAMD_F64)
      hipLaunchKernelGGL(k_gauge_xfer<double>, grid, dim3(256), 0, stream(), u->d, (double *)st, g, mu, m, to_host, 0);
    else
      hipLaunchKernelGGL(k_gauge_xfer<float>, grid, dim3(256), 0, stream(), u->d, (float *)st, g, mu, m, to_host, 0);
    QAMD_HIP(hipGetLastError());
    if (to_host && d2h_pageable(host[mu], st, bytes)) return -1;
  }
  return 0;
}

int qamd_gauge_from_qll(qamd_gauge *u, const qamd_qll_layout *l, const void *const host[4], int hprec) {
  return gauge_xfer_qll(u, l, (void *const *)host, hprec, 0);
}
int qamd_gauge_to_qll(const qamd_gauge *u, const qamd_qll_layout *l, void *const host[4], int hprec) {
  return gauge_xfer_qll((qamd_gauge *)u, l, host, hprec, 1);
}

static int gauge_xfer_sites(qamd_gauge *u, const int *coords, int64_t nsites, void *const host[4], int hprec,
                            int row_major, int to_host) {
  const Geom &g = u->lat->g;
  if (hprec != QAMD_F64 && hprec != QAMD_F32) {
    set_error("gauge site-table transfer: host precision must be QAMD_F64 or QAMD_F32");
    return -1;
  }
  SrcMap m;
  memset(&m, 0, sizeof m);
  m.kind = 2;
  int *inv = nullptr;
  if (site_table_inverse(u->lat, coords, nsites, 2, &inv)) return -1;
  m.inv = inv;
  size_t bytes = (size_t)nsites * 18 * (hprec == QAMD_F64 ? 8 : 4);
  void *st = dev_alloc(bytes);
  if (!st) {
    dev_free(inv);
    return -1;
  }
  hipError_t e = hipSuccess;
  for (int mu = 0; mu < 4 && e == hipSuccess; mu++) {
    if (!to_host) e = hipMemcpyAsync(st, host[mu], bytes, hipMemcpyHostToDevice, stream());
    dim3 grid((unsigned)((g.vh + 255) / 256), 2);
    if (hprec == QAMD_F64)
      hipLaunchKernelGGL(k_gauge_xfer<double>, grid, dim3(256), 0, stream(), u->d, (double *)st, g, mu, m, to_host,
                         row_major);
    else
      hipLaunchKernelGGL(k_gauge_xfer<float>, grid, dim3(256), 0, stream(), u->d, (float *)st, g, mu, m, to_host,
                         row_major);
    if (e == hipSuccess) e = hipGetLastError();
    if (to_host && e == hipSuccess) e = hipMemcpyAsync(host[mu], st, bytes, hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
  }
  dev_free(st);
  dev_free(inv);
  if (e != hipSuccess) {
    set_error("gauge site-table transfer: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int qamd_gauge_from_sites(qamd_gauge *u, const int *coords, int64_t nsites, const void *const host[4], int hprec,
                          int row_major) {
  return gauge_xfer_sites(u, coords, nsites, (void *const *)host, hprec, row_major, 0);
}

int qamd_gauge_to_sites(const qamd_gauge *u, const int *coords, int64_t nsites, void *const host[4], int hprec,
                        int row_major) {
  return gauge_xfer_sites((qamd_gauge *)u, coords, nsites, host, hprec, row_major, 1);
}

int qamd_gauge_random_su3(qamd_gauge *u, uint64_t seed) {
  hipLaunchKernelGGL(k_random_su3, gauge_grid(u->lat->g, 256), dim3(256), 0, stream(), u->d, u->lat->g, seed);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int qamd_gauge_stag_phases(qamd_gauge *u) {
  hipLaunchKernelGGL(k_stag_phases, gauge_grid(u->lat->g, 256), dim3(256), 0, stream(), u->d, u->lat->g);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int qamd_fat7(qamd_gauge *fl, qamd_gauge *ll, const double coef[6], const qamd_gauge *u,
              const qamd_gauge *ulong) {
  if (!fl || !u) { set_error("fat7: NULL field"); return -1; }
  if (fl == u || (ll && (ll == u || ll == ulong))) { set_error("fat7: outputs must not alias inputs"); return -1; }
  const Geom &g = u->lat->g;
  Fat7Coef C = {coef[0], coef[1], coef[2], coef[3], coef[4], coef[5]};
  hipLaunchKernelGGL(k_fat7, gauge_grid(g, 128), dim3(128), 0, stream(), fl->d, u->d, g, C);
  QAMD_HIP(hipGetLastError());
  if (ll) {
    const qamd_gauge *ul = ulong ? ulong : u;
    hipLaunchKernelGGL(k_longlink, gauge_grid(g, 256), dim3(256), 0, stream(), ll->d, ul->d, g, coef[5]);
    QAMD_HIP(hipGetLastError());
  }
  return 0;
}

int qamd_gauge_plaquette(const qamd_gauge *u, double plaq[6]) {
  const Geom &g = u->lat->g;
  int nb = 1024;
  if (red_reserve(nb)) return -1;
  double *slot = scalar_slot();
  if (!slot) return -1;
  FinStore6 fin{slot, 1.0 / (3.0 * (double)g.V)};
  RedScratch rs = red_set(0);
  hipLaunchKernelGGL(k_plaq, dim3(nb), dim3(256), 0, stream(), u->d, g, rs);
  launch_finish<6>(rs, nb, fin, stream());
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipMemcpyAsync(plaq, slot, 6 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  return 0;
}

}  // extern "C"
