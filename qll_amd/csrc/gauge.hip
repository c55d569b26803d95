// Gauge fields on the device: seeded SU(3) generation, staggered phases,
// fat7 + Naik smearing in three staple passes per direction, plaquette,
// host transfer.
//
// fat7 (lib/gauge.c:188-416) in FP64, with
//   St_nu[X](x) = U_nu(x) X(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ X(x-nu) U_nu(x-nu+mu)
//   fat_mu = (c1 - 6 cL) U_mu + c3 sum_nu St_nu[U_mu] + cL sum_nu St_nu[St_nu[U_mu]]
//          + c5 sum_{nu,rho} St_rho[St_nu[U_mu]]
//          + c7 sum_{nu,rho,sig} St_sig[St_rho[St_nu[U_mu]]]
// the sum QLL's genStaple recursion builds with transporters and temporary
// fields.
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
// A += a B C (nn) and A += a B C^+ (na), without a temporary for the product
__device__ inline void madd_nn(M3 &A, double a, const M3 &B, const M3 &C) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double2 t = cmul(B.e[3 * k + r], C.e[3 * c + k]);
        s.x += t.x;
        s.y += t.y;
      }
      A.e[3 * c + r].x += a * s.x;
      A.e[3 * c + r].y += a * s.y;
    }
}
__device__ inline void madd_na(M3 &A, double a, const M3 &B, const M3 &C) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double2 t = cmulc(B.e[3 * k + r], C.e[3 * k + c]);
        s.x += t.x;
        s.y += t.y;
      }
      A.e[3 * c + r].x += a * s.x;
      A.e[3 * c + r].y += a * s.y;
    }
}
__device__ inline void madd(M3 &A, double a, const M3 &B) {
#pragma unroll
  for (int i = 0; i < 9; i++) {
    A.e[i].x += a * B.e[i].x;
    A.e[i].y += a * B.e[i].y;
  }
}

// site coordinates as four named values: a direction known only at run time
// selects among them (no indexed array, which would live in scratch memory)
struct SC {
  int c0, c1, c2, c3;
};
__device__ __forceinline__ SC sc_of(const Site &s) { return SC{s.c[0], s.c[1], s.c[2], s.c[3]}; }

struct GView {
  const double2 *d;
  Geom g;
  __device__ M3 load(int f, const SC &c) const {   // field f (a direction) at site c
    const int lex = c.c0 + g.stride[1] * c.c1 + g.stride[2] * c.c2 + g.stride[3] * c.c3;
    const int par = (c.c0 + c.c1 + c.c2 + c.c3) & 1;
    const double2 *p = d + ((int64_t)(f * 2 + par) * 9) * g.vh_pad + (lex >> 1);
    M3 m;
#pragma unroll
    for (int e = 0; e < 9; e++) m.e[e] = p[e * g.vh_pad];
    return m;
  }
  __device__ SC shifted(SC c, int mu, int h) const {   // c + h mu (periodic)
    const int L = sel4(mu, g.L[0], g.L[1], g.L[2], g.L[3]);
    int v = sel4(mu, c.c0, c.c1, c.c2, c.c3) + h;
    v += (v < 0) ? L : 0;
    v -= (v >= L) ? L : 0;
    c.c0 = mu == 0 ? v : c.c0;
    c.c1 = mu == 1 ? v : c.c1;
    c.c2 = mu == 2 ? v : c.c2;
    c.c3 = mu == 3 ? v : c.c3;
    return c;
  }
};

struct Fat7Coef {
  double c1, c3, c5, c7, cL, naik;
};

// ------------------------------------------- fat7 as three field passes
// The staples are built once per (site, direction) and shared, as QLL's
// genStaple sequence does (lib/gauge.c:264-366), instead of recomputed along
// every path: for each direction mu,
//   pass 1: S3_nu = St_nu[U_mu] for the three nu != mu (stored),
//           fat_mu = (c1 - 6 cL) U_mu + c3 sum_nu S3_nu
//   pass 2: T5_sig = sum over the two orderings (nu, rho) of the directions
//           other than mu and sig of St_rho[S3_nu] (stored),
//   pass 3: fat_mu += cL sum_nu St_nu[S3_nu] + c5 sum_sig T5_sig
//                   + c7 sum_sig St_sig[T5_sig]
// (each 7-staple St_sig[St_rho[St_nu[U]]] has sig = the one direction left,
// so the staple is linear in its argument and the two 5-staples that feed
// one sig share it).  Temporaries: 12 + 12 matrix fields in the gauge
// layout, field 3 mu + j for the j-th direction other than mu.
//
// Thread mapping: a 256-thread block is 64 sites x the 4 directions (wave w
// computes direction w), blocks alternate parity over the same site range,
// and consecutive blocks run on one die (XCD order), so the four directions'
// loads of a site and the neighbours of the other parity meet in one L2.
// Each pass holds one accumulator (two would not fit the 256 registers of
// two waves per SIMD: the staple's six operand loads are issued together).

// the j-th direction other than mu, and the inverse
__device__ __forceinline__ int fat_dir(int mu, int j) { return j + (j >= mu); }
__device__ __forceinline__ int fat_idx(int mu, int nu) { return nu - (nu > mu); }

// St_nu[X](x) = U_nu(x) X(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ X(x-nu) U_nu(x-nu+mu)
// for the mu-direction field fx of view X
__device__ M3 fat_staple(const GView &U, const GView &X, int fx, int mu, int nu, const SC &c) {
  M3 t = mul_nn(U.load(nu, c), X.load(fx, U.shifted(c, nu, 1)));
  M3 res = mul_na(t, U.load(nu, U.shifted(c, mu, 1)));
  const SC e = U.shifted(c, nu, -1);
  M3 t2 = mul_an(U.load(nu, e), X.load(fx, e));
  madd(res, 1.0, mul_nn(t2, U.load(nu, U.shifted(e, mu, 1))));
  return res;
}
// A += a St_nu[X](x), the staple not held as a matrix
__device__ void fat_staple_acc(M3 &A, double a, const GView &U, const GView &X, int fx, int mu, int nu,
                               const SC &c) {
  {
    const M3 t = mul_nn(U.load(nu, c), X.load(fx, U.shifted(c, nu, 1)));
    madd_na(A, a, t, U.load(nu, U.shifted(c, mu, 1)));
  }
  const SC e = U.shifted(c, nu, -1);
  const M3 t2 = mul_an(U.load(nu, e), X.load(fx, e));
  madd_nn(A, a, t2, U.load(nu, U.shifted(e, mu, 1)));
}

// Outputs are stored non-temporally: a pass writes 1-3 KB per site, which
// would otherwise take the die's L2 from the gauge field and staples the
// neighbouring sites are still to read.
__device__ __forceinline__ void nt_store2(double2 *p, const double2 &v) {
  __builtin_nontemporal_store(dvec2_t{v.x, v.y}, reinterpret_cast<dvec2_t *>(p));
}
__device__ __forceinline__ void fat_store(double2 *base, const Geom &g, int f, int par, int cb, const M3 &m) {
  double2 *p = base + ((int64_t)(f * 2 + par) * 9) * g.vh_pad + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) nt_store2(p + e * g.vh_pad, m.e[e]);
}
__device__ __forceinline__ void fat_accum(double2 *base, const Geom &g, int f, int par, int cb, const M3 &m) {
  double2 *p = base + ((int64_t)(f * 2 + par) * 9) * g.vh_pad + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const dvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const dvec2_t *>(p + e * g.vh_pad));
    nt_store2(p + e * g.vh_pad, make_double2(v.x + m.e[e].x, v.y + m.e[e].y));
  }
}

struct FatSite {
  int mu, par, cb, k;
};
// nk: blocks per 64-site run (k = 0..nk-1, consecutive, so they share a die).
// so.P > 0: the runs in slab-sweep order (die x walks its piece of every
// t-slice, slab_block over 64-site runs), else XCD-contiguous ranges
__device__ __forceinline__ FatSite fat_site(const SlabOrder &so, int nk = 1) {
  if (so.P > 0) {
    const int x = blockIdx.x % NXCD, j = blockIdx.x / NXCD, sb = j / nk;
    const int run = slab_block((sb >> 1) * NXCD + x, so);
    return FatSite{(int)(threadIdx.x >> 6), sb & 1, run * 64 + (int)(threadIdx.x & 63), j - sb * nk};
  }
  const int b = xcd_block(blockIdx.x, gridDim.x), sb = b / nk;
  return FatSite{(int)(threadIdx.x >> 6), sb & 1, (sb >> 1) * 64 + (int)(threadIdx.x & 63), b - sb * nk};
}
static dim3 fat_grid(const Geom &g, int nk = 1) { return dim3((unsigned)(2 * nk * ((g.vh + 63) / 64))); }

// also the long links (ll != NULL): naik ul_mu(x) ul_mu(x+mu) ul_mu(x+2mu)
// (lib/gauge.c:369-392), whose loads share the block's U traffic when ul = u
__global__ void __launch_bounds__(256) k_fat7_p1(double2 *fl, double2 *s3, double2 *ll, const double2 *u,
                                                 const double2 *ul, Geom g, Fat7Coef C, int keep, SlabOrder so) {
  const FatSite f = fat_site(so);
  if (f.cb >= g.vh) return;
  const SC c = sc_of(site_of(g, f.cb, f.par));
  GView U{u, g};
  if (ll) {
    GView L{ul, g};
    const SC c1 = L.shifted(c, f.mu, 1);
    M3 A = mul_nn(L.load(f.mu, c), L.load(f.mu, c1));
    A = mul_nn(A, L.load(f.mu, L.shifted(c1, f.mu, 1)));
#pragma unroll
    for (int i = 0; i < 9; i++) {
      A.e[i].x *= C.naik;
      A.e[i].y *= C.naik;
    }
    fat_store(ll, g, f.mu, f.par, f.cb, A);
  }
  M3 F = U.load(f.mu, c);
  const double c1 = C.c1 - 6 * C.cL;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    F.e[i].x *= c1;
    F.e[i].y *= c1;
  }
#pragma unroll 1
  for (int j = 0; j < 3; j++) {
    M3 S = fat_staple(U, U, f.mu, f.mu, fat_dir(f.mu, j), c);
    if (keep) fat_store(s3, g, 3 * f.mu + j, f.par, f.cb, S);
    madd(F, C.c3, S);
  }
  fat_store(fl, g, f.mu, f.par, f.cb, F);
}

// three blocks per site run: T5 for sig = the k-th direction other than mu
__global__ void __launch_bounds__(256) k_fat7_p2(double2 *t5, const double2 *s3, const double2 *u, Geom g,
                                                 SlabOrder so) {
  const FatSite f = fat_site(so, 3);
  if (f.cb >= g.vh) return;
  const int mu = f.mu, k = f.k, sig = fat_dir(mu, k);
  const SC c = sc_of(site_of(g, f.cb, f.par));
  GView U{u, g}, S3{s3, g};
  M3 T;
#pragma unroll
  for (int i = 0; i < 9; i++) T.e[i] = make_double2(0.0, 0.0);
#pragma unroll 1
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu || nu == sig) continue;
    const int rho = 6 - mu - sig - nu;   // the fourth direction
    fat_staple_acc(T, 1.0, U, S3, 3 * mu + fat_idx(mu, nu), mu, rho, c);
  }
  fat_store(t5, g, 3 * mu + k, f.par, f.cb, T);
}

// fat_mu += cL sum_nu St_nu[S3_nu] + c5 sum_sig T5_sig + c7 sum_sig St_sig[T5_sig]
__global__ void __launch_bounds__(256) k_fat7_p3(double2 *fl, const double2 *s3, const double2 *t5, const double2 *u,
                                                 Geom g, Fat7Coef C, SlabOrder so) {
  const FatSite f = fat_site(so);
  if (f.cb >= g.vh) return;
  const int mu = f.mu;
  const SC c = sc_of(site_of(g, f.cb, f.par));
  GView U{u, g}, S3{s3, g}, T5{t5, g};
  M3 F;
#pragma unroll
  for (int i = 0; i < 9; i++) F.e[i] = make_double2(0.0, 0.0);
  if (C.cL != 0)
#pragma unroll 1
    for (int j = 0; j < 3; j++) fat_staple_acc(F, C.cL, U, S3, 3 * mu + j, mu, fat_dir(mu, j), c);
  if (C.c5 != 0 || C.c7 != 0)
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
      if (C.c5 != 0) madd(F, C.c5, T5.load(3 * mu + k, c));
      if (C.c7 != 0) fat_staple_acc(F, C.c7, U, T5, 3 * mu + k, mu, fat_dir(mu, k), c);
    }
  fat_accum(fl, g, mu, f.par, f.cb, F);
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
    const SC c = sc_of(site_of(g, cb, par));
    int l = 0;
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
      for (int nu = 0; nu < mu; nu++) {
        M3 A = mul_nn(U.load(mu, c), U.load(nu, U.shifted(c, mu, 1)));
        M3 B = mul_nn(U.load(nu, c), U.load(mu, U.shifted(c, nu, 1)));
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

// fewest sub-pieces (y splits before z splits) that bring a die's piece to
// at most `sites` sites (both parities) per t-slice; the finest split when
// none does
static void fat_split(SlabOrder &s, int sites) {
  int by = 1, bz = 1;
  for (int n = 1; n <= s.bpy * s.zs; n++) {
    for (int ny = n; ny >= 1; ny--) {
      if (n % ny != 0) continue;
      const int nz = n / ny;
      if (s.bpy % ny != 0 || s.zs % nz != 0) continue;
      by = ny;
      bz = nz;
      if ((int64_t)2 * 64 * (s.bpy / ny) * (s.zs / nz) <= sites) {
        (void)slab_split(s, ny, nz);
        return;
      }
      break;
    }
  }
  (void)slab_split(s, by, bz);
}

int qamd_fat7(qamd_gauge *fl, qamd_gauge *ll, const double coef[6], const qamd_gauge *u,
              const qamd_gauge *ulong) {
  if (!fl || !u) { set_error("fat7: NULL field"); return -1; }
  if (fl == u || (ll && (ll == u || ll == ulong))) { set_error("fat7: outputs must not alias inputs"); return -1; }
  const Geom &g = u->lat->g;
  Fat7Coef C = {coef[0], coef[1], coef[2], coef[3], coef[4], coef[5]};
  {
    const bool have5 = C.c5 != 0 || C.c7 != 0 || C.cL != 0;   // S3 kept for the nested staples
    const bool have57 = C.c5 != 0 || C.c7 != 0;                 // the T5 sums needed
    const size_t nf = (size_t)2 * 9 * g.vh_pad;   // double2 per matrix field
    double2 *tmp = (double2 *)dev_alloc(((have5 ? 12 : 0) + (have57 ? 12 : 0)) * nf * sizeof(double2) + 256);
    if (!tmp) return -1;
    double2 *s3 = tmp, *t5 = tmp + (have5 ? 12 : 0) * nf;
    const dim3 grid = fat_grid(g);
    // runs of 64 sites in the slab-sweep order where the lattice splits that
    // way (32^4: 9 % less fetched, 5 % faster than XCD-contiguous ranges,
    // which QAMD_FAT7_ORDER=0 selects); same sums either way
    SlabOrder so;
    memset(&so, 0, sizeof so);
    const char *fo = getenv("QAMD_FAT7_ORDER");
    if ((fo && atoi(fo) == 0) || !slab_order(g, so, 64)) memset(&so, 0, sizeof so);
    // passes 2 and 3 read 12-24 staple fields per site besides U, so their
    // t-window outgrows a die's L2 first: they sweep each die's piece in
    // sub-pieces of at most 1024 (pass 2) and 2048 (pass 3) sites per t-slice
    // (32^4: pieces of 4096 split 4 and 2 ways in y, 2-4 % faster per pass
    // than whole pieces; pass 2 split 2 ways runs within 1 % of 4 ways but
    // fetches 8 % more; pass 1 is fastest whole).
    // QAMD_FAT7_SUB=ny1,nz1,ny2,nz2,ny3,nz3 sets the splits of the three
    // passes instead.
    SlabOrder sp[3] = {so, so, so};
    if (so.P > 0) {
      const char *fs = getenv("QAMD_FAT7_SUB");
      int sub[6] = {1, 1, 1, 1, 1, 1};
      if (fs && sscanf(fs, "%d,%d,%d,%d,%d,%d", &sub[0], &sub[1], &sub[2], &sub[3], &sub[4], &sub[5]) == 6) {
        for (int k = 0; k < 3; k++) (void)slab_split(sp[k], sub[2 * k], sub[2 * k + 1]);
      } else {
        fat_split(sp[1], 1024);
        fat_split(sp[2], 2048);
      }
    }
    hipLaunchKernelGGL(k_fat7_p1, grid, dim3(256), 0, stream(), fl->d, s3, ll ? ll->d : nullptr, u->d,
                       ulong ? ulong->d : u->d, g, C, have5 ? 1 : 0, sp[0]);
    if (have57) hipLaunchKernelGGL(k_fat7_p2, fat_grid(g, 3), dim3(256), 0, stream(), t5, s3, u->d, g, sp[1]);
    if (have5) hipLaunchKernelGGL(k_fat7_p3, grid, dim3(256), 0, stream(), fl->d, s3, t5, u->d, g, C, sp[2]);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(stream());
    dev_free(tmp);
    if (e != hipSuccess) {
      set_error("fat7: %s", hipGetErrorString(e));
      return -1;
    }
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
