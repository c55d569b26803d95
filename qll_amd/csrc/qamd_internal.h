// Internal definitions shared by the HIP translation units of libqll_amd.
//
// Device data layout (HBM), per lattice parity p in {0 = even, 1 = odd}:
//   * a site of parity p is addressed by its checkerboard index
//       cb = lex >> 1,  lex = x + Lx*(y + Ly*(z + Lz*t))
//     (all extents even), so the neighbour of a site is found by index
//     arithmetic, with no tables;
//   * colour vectors / spinors are structure-of-arrays: complex component j
//     of rhs r at site cb lives at [(r*ncomp + j) * vh_pad + cb];  FP16 fields
//     keep one FP32 max-norm per (rhs, site) beside them;
//   * link sets are blocked by 64 sites (one wavefront) and packed so that
//     one 16-byte load returns element k of G = 16/sizeof(complex) consecutive
//     storage slots (G = 1 FP64, 2 FP32, 4 FP16/Q16): element k of slot sl
//     for site cb lives at [(((cb>>6)*(nterm/G)*9 + row)*64 + (cb&63))*G +
//     sl%G] with row = link_row<G>(sl, k, nterm) (below), so a wavefront
//     streams one contiguous block per site group with full-width loads.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stddef.h>
#include <utility>

#include "qll_amd.h"

namespace qamd {

constexpr int WAVE = 64;
constexpr int NXCD = 8;   // MI355X: 8 accelerator complex dies, blocks round-robin

// ------------------------------------------------------------ error handling
void set_error(const char *fmt, ...);
#define QAMD_HIP(call)                                                        \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::qamd::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,          \
                        hipGetErrorString(_e));                               \
      return -1;                                                              \
    }                                                                         \
  } while (0)

hipStream_t stream();
int ensure_init();
// Capture what enqueue(ctx) launches on the library stream into an
// executable graph (kernel launches only).  On any failure the stream leaves
// capture mode, partial objects are destroyed and NULL is returned, so the
// caller can launch eagerly; QAMD_TEST_FAIL_CAPTURE=1 forces that failure.
hipGraphExec_t capture_stream(int (*enqueue)(void *), void *ctx);
bool initialized();   // qamd_init (or a first library call) has run
void *dev_alloc(size_t bytes);   // returns NULL and sets the error on failure
void dev_free(void *p);
// transfers of pageable host arrays (hostio.hip): a persistent device staging
// buffer of at least `bytes` (valid until the next stage_dev call), and
// chunked copies through pinned buffers on the library stream.  h2d returns
// once the host data has been read (the DMA may still be in flight, ordered
// on the stream); d2h returns with the host data complete.
void *stage_dev(size_t bytes);
int h2d_pageable(void *dst_dev, const void *src_host, size_t bytes);
int d2h_pageable(void *dst_host, const void *src_dev, size_t bytes);

// ---------------------------------------------- fast division by invariants
// Quotient of a non-negative 31-bit integer by a fixed divisor with one
// multiply-high and one shift (Granlund & Montgomery, PLDI 1994).
struct FastDiv {
  uint32_t d, mul, shr;
  __host__ void init(uint32_t div) {
    d = div;
    uint32_t p = 0;
    while ((1u << p) < div) p++;
    if (div == 1) { mul = 0; shr = 0; }
    else {
      mul = (uint32_t)((((uint64_t)1 << (31 + p)) + div - 1) / div);
      shr = p - 1;
    }
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return d == 1 ? n : (__umulhi(n, mul) >> shr);
#else
    return d == 1 ? n : (uint32_t)(((uint64_t)n * mul) >> 32) >> shr;
#endif
  }
};

// ---------------------------------------------------------------- geometry
struct Geom {
  int L[4];
  int64_t V, vh, vh_pad;   // volume, sites per parity, padded to 64
  int stride[4];           // lexicographic strides
  FastDiv dxh, dy, dz;     // divisors Lx/2, Ly, Lz
};
void geom_init(Geom &g, const int L[4]);

struct Site {
  int c[4];
  int lex;
};

// coordinates of checkerboard site cb of parity par
__device__ __forceinline__ Site site_of(const Geom &g, int cb, int par) {
  Site s;
  uint32_t q1 = g.dxh.div((uint32_t)cb);
  int xh = cb - (int)q1 * (g.L[0] >> 1);
  uint32_t q2 = g.dy.div(q1);
  s.c[1] = (int)q1 - (int)q2 * g.L[1];
  uint32_t q3 = g.dz.div(q2);
  s.c[2] = (int)q2 - (int)q3 * g.L[2];
  s.c[3] = (int)q3;
  int xb = (s.c[1] + s.c[2] + s.c[3] + par) & 1;
  s.c[0] = 2 * xh + xb;
  s.lex = 2 * cb + xb;
  return s;
}

// Tiled site order: logical site i of parity par walks tiles of
// 2^s[0] x 2^s[1] x 2^s[2] x 2^s[3] sites (s[0] >= 1), half of them of each
// parity, x fastest inside a tile and tiles x fastest.  Keeps the neighbours
// of a sweep inside a small window of recently read sites.
struct Tiling {
  int s[4];            // log2 tile extents
  FastDiv nx, ny, nz;  // tiles per direction
  __host__ bool init(const int L[4], const int sh[4]) {
    if (sh[0] < 1) return false;
    for (int i = 0; i < 4; i++) {
      s[i] = sh[i];
      if ((1 << sh[i]) > L[i] || L[i] % (1 << sh[i])) return false;
    }
    nx.init(L[0] >> s[0]);
    ny.init(L[1] >> s[1]);
    nz.init(L[2] >> s[2]);
    return true;
  }
};

__host__ __device__ __forceinline__ Site site_of_tiled(const Geom &g, const Tiling &T, int i, int par) {
  const int bx = T.s[0] - 1, per = bx + T.s[1] + T.s[2] + T.s[3];   // log2 sites per parity per tile
  const uint32_t tile = (uint32_t)i >> per;
  int j = i & ((1 << per) - 1);
  const int xh = j & ((1 << bx) - 1);
  j >>= bx;
  const int y = j & ((1 << T.s[1]) - 1);
  j >>= T.s[1];
  const int z = j & ((1 << T.s[2]) - 1);
  const int t = j >> T.s[2];
  const uint32_t q1 = T.nx.div(tile), q2 = T.ny.div(q1), q3 = T.nz.div(q2);
  Site s;
  s.c[1] = (((int)q1 - (int)q2 * (int)T.ny.d) << T.s[1]) + y;
  s.c[2] = (((int)q2 - (int)q3 * (int)T.nz.d) << T.s[2]) + z;
  s.c[3] = ((int)q3 << T.s[3]) + t;
  const int xb = (s.c[1] + s.c[2] + s.c[3] + par) & 1;
  s.c[0] = (((int)tile - (int)q1 * (int)T.nx.d) << T.s[0]) + 2 * xh + xb;
  s.lex = s.c[0] + s.c[1] * g.stride[1] + s.c[2] * g.stride[2] + s.c[3] * g.stride[3];
  return s;
}

// checkerboard index of the neighbour at hop h (|h| <= 3 < L) along mu
__device__ __forceinline__ int nbr_cb(const Geom &g, const Site &s, int mu, int h) {
  int c = s.c[mu] + h;
  int L = g.L[mu];
  c += (c < 0) ? L : 0;
  c -= (c >= L) ? L : 0;
  return (s.lex + (c - s.c[mu]) * g.stride[mu]) >> 1;
}

// link packing: terms per 16-byte load for stored complex type S
template <class S> struct LinkG { static constexpr int G = (int)(16 / sizeof(S)); };

// the same for a direction known only at run time (selects, no indexed
// access to the coordinate array)
template <class T>
__device__ __forceinline__ T sel4(int mu, T a, T b, T c, T d) {   // values, not lvalues
  return mu == 0 ? a : mu == 1 ? b : mu == 2 ? c : d;
}
__device__ __forceinline__ int nbr_cb_rt(const Geom &g, const Site &s, int mu, int h) {
  const int c = sel4(mu, s.c[0], s.c[1], s.c[2], s.c[3]);
  const int L = sel4(mu, g.L[0], g.L[1], g.L[2], g.L[3]);
  const int st = sel4(mu, g.stride[0], g.stride[1], g.stride[2], g.stride[3]);
  int n = c + h;
  n += (n < 0) ? L : 0;
  n -= (n >= L) ? L : 0;
  return (s.lex + (n - c) * st) >> 1;
}

// Storage slot of term t.  In a 16-term staggered set the 1-hop terms
// (t = 2k) fill slots 0-7 and the 3-hop terms (t = 2k+1) slots 8-15, so a
// 16-byte group of G slots never mixes the two kinds and a long-term group can
// skip its third-row loads as a whole.  8-term sets store term t in slot t.
__host__ __device__ constexpr int term_slot(int t, int nterm) {
  return nterm == 16 ? ((t & 1) << 3) | (t >> 1) : t;
}
__host__ __device__ constexpr int slot_term(int sl, int nterm) {
  return nterm == 16 ? ((sl & 7) << 1) | (sl >> 3) : sl;
}

// Processing order of a site's slot groups: a 1-hop group, then the 3-hop
// group of the same directions, and so on (QLL's interleaved term order at
// group granularity), so that the gathers of one direction's 1- and 3-hop
// neighbours stay close in time.  8-term sets: storage order.
template <int G, int NT>
__host__ __device__ constexpr int group_at(int o) {
  return NT == 16 ? ((o & 1) ? (8 / G) + (o >> 1) : (o >> 1)) : o;
}

// Row (a run of 64 sites x G slots) of element k of slot sl inside one
// 64-site block.  The eight terms whose third row a Dslash may rebuild (the
// 3-hop slots 8-15 of a 16-term staggered set; all eight terms of an 8-term
// set) store the first two rows (k%3 != 2) of every group, then all third
// rows; a 16-term set puts its 1-hop groups (9 rows each) in front.  A Dslash
// that rebuilds the third rows then reads one contiguous prefix of every
// block (16 terms: 120 of 144 rows; 8 terms: 48 of 72).
template <int G>
__host__ __device__ constexpr int link_row(int sl, int k, int nterm) {
  if (nterm == 16 && sl < 8) return (sl / G) * 9 + k;
  const int base = nterm == 16 ? (8 / G) * 9 : 0, g = (nterm == 16 ? sl - 8 : sl) / G;
  return (k % 3 != 2) ? base + g * 6 + (k / 3) * 2 + k % 3 : base + (8 / G) * 6 + g * 3 + k / 3;
}

// Forward-only (8-slot) sets keep each slot's elements together instead, so
// that a Dslash reading a single slot of a neighbour's set (a backward term)
// uses every byte of its 16-byte words.  Per 64-site block (rows of 64 sites
// x 16 bytes):
//   G = 1: the 8-term layout above;
//   G = 2: slot sl in rows 4sl..4sl+3 as the words {0,1} {3,4} {6,7} {2,5},
//          element 8 of slots 2j, 2j+1 in row 32 + j;
//   G = 4: slot sl in rows 2sl, 2sl+1 as {0,1,3,4} {6,7,2,5}, element 8 of
//          slots 4j..4j+3 in row 16 + j.
// The first two rows of a term (elements 0,1,3,4,6,7) are the leading
// elements of its words.
template <int G>
__host__ __device__ constexpr int fb_row(int sl, int k) {
  if (G == 1) return link_row<1>(sl, k, 8);
  if (k == 8) return 8 * (8 / G) + sl / G;
  if (G == 2) return sl * 4 + (k % 3 == 2 ? 3 : k / 3);
  return sl * 2 + ((k % 3 == 2 || k >= 6) ? 1 : 0);
}
template <int G>
__host__ __device__ constexpr int fb_pos(int sl, int k) {
  if (G == 1) return 0;
  if (k == 8) return sl % G;
  if (G == 2) return (k % 3 == 2) ? k / 3 : k % 3;
  return (k % 3 == 2) ? 2 + k / 3 : (k >= 6 ? k - 6 : k % 3 + 2 * (k / 3));
}

// 16-term staggered sets of FP32 and 16-bit links keep their 1-hop terms
// (slots 0-7) and their 3-hop terms (slots 8-15) in two regions of the
// array, each in the block layout above with (8/G)*9 rows per 64-site block:
// region 1 starts 72*vh_pad complex elements in (8 slots x 9 elements).  A
// Dslash wavefront then reads two dense streams instead of one interleaved
// one; on MI355X at 32^4 that measured 1-3 % faster (FP32 single-rhs Dslash
// 0.094 -> 0.093 ms, FP32 8-rhs block 0.180 -> 0.176 ms, Q16 8-rhs block
// 0.156 -> 0.152 ms).  FP64 sets measured no gain and keep one region.
template <class S>
__host__ __device__ constexpr bool link_split(int nterm) { return nterm == 16 && LinkG<S>::G >= 2; }

// blocked link index: element k (0..8) of term t for site cb (fb: the
// forward-only layout, nterm = 8); vp = sites per parity padded to 64
template <class S>
__host__ __device__ __forceinline__ int64_t link_idx(int64_t cb, int t, int k, int nterm, bool fb, int64_t vp) {
  constexpr int G = LinkG<S>::G;
  if (fb) return ((((cb >> 6) * (8 / G) * 9 + fb_row<G>(t, k)) << 6) + (cb & 63)) * G + fb_pos<G>(t, k);
  const int sl = term_slot(t, nterm);
  if (link_split<S>(nterm)) {
    constexpr int RB = (8 / G) * 9;   // rows per 64-site block in each region
    const int r = link_row<G>(sl, k, 16);
    const bool hop3 = sl >= 8;
    return (hop3 ? (int64_t)72 * vp : 0) + ((((cb >> 6) * RB + (hop3 ? r - RB : r)) << 6) + (cb & 63)) * G + sl % G;
  }
  return ((((cb >> 6) * (nterm / G) * 9 + link_row<G>(sl, k, nterm)) << 6) + (cb & 63)) * G + sl % G;
}
// offset of (slot, k) from link_base<S, 8> in the forward-only layout
template <class S>
__device__ __forceinline__ int link_off_fb(int sl, int k) {
  constexpr int G = LinkG<S>::G;
  return fb_row<G>(sl, k) * 64 * G + fb_pos<G>(sl, k);
}
// first element of site cb's link stream (of region 0 for split sets), the
// first of its 3-hop region (the same pointer for sets in one region), and
// the offset of (slot, k) from the base of the slot's region
template <class S, int NT>
__device__ __forceinline__ const S *link_base(const S *links, int cb) {
  constexpr int G = LinkG<S>::G, RB = link_split<S>(NT) ? (8 / G) * 9 : NT * 9 / G;
  return links + (int64_t)(cb >> 6) * RB * 64 * G + (int64_t)(cb & 63) * G;
}
template <class S, int NT>
__device__ __forceinline__ const S *link_base3(const S *lp, int64_t vp) {
  return link_split<S>(NT) ? lp + 72 * vp : lp;
}
template <class S, int NT>
__device__ __forceinline__ int link_off(int sl, int k) {
  constexpr int G = LinkG<S>::G, RB = (8 / G) * 9;
  const int r = link_row<G>(sl, k, NT);
  return ((link_split<S>(NT) && sl >= 8) ? r - RB : r) * 64 * G + sl % G;
}

// Third row of a term c*W (W unitary, det W = -1 if neg) from its first
// two: det W * conj(row0 x row1) / c, c = |row0|.  Element (r, c) is m[3c+r].
// Scale-free, so it also holds for 16-bit terms read as raw integers.
// Written as explicit FMA chains with contraction off: which products the
// compiler would fuse on its own depends on the surrounding code, and every
// kernel instance (rhs count, reduction, launch shape) must rebuild the same
// bits (2 products + 6 FMAs per component instead of 8 products + 6 adds).
template <class C, class R>
__device__ __forceinline__ void long_row3(C (&m)[9], bool neg) {
#pragma clang fp contract(off)
  R c2 = m[0].x * m[0].x;
  c2 = fma(m[0].y, m[0].y, c2);
  c2 = fma(m[3].x, m[3].x, c2);
  c2 = fma(m[3].y, m[3].y, c2);
  c2 = fma(m[6].x, m[6].x, c2);
  c2 = fma(m[6].y, m[6].y, c2);
  const R ic = (neg ? (R)-1 : (R)1) / sqrt(c2);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const C a1 = m[3 * ((c + 1) % 3)], a2 = m[3 * ((c + 2) % 3)];
    const C b1 = m[3 * ((c + 1) % 3) + 1], b2 = m[3 * ((c + 2) % 3) + 1];
    // (a1 b2 - a2 b1) componentwise in complex arithmetic
    R re = a1.x * b2.x;
    re = fma(-a1.y, b2.y, re);
    re = fma(-a2.x, b1.x, re);
    re = fma(a2.y, b1.y, re);
    R im = a1.x * b2.y;
    im = fma(a1.y, b2.x, im);
    im = fma(-a2.x, b1.y, im);
    im = fma(-a2.y, b1.x, im);
    m[3 * c + 2].x = re * ic;
    m[3 * c + 2].y = -im * ic;
  }
}

// XCD-aware block order: block b runs on die b % 8; give each die one
// contiguous range of logical blocks so that neighbouring sites share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int xcd = b % NXCD, idx = b / NXCD;
  int q = nb / NXCD, r = nb % NXCD;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

// Slab-sweep block order for 256-site blocks of one parity.  Logical block
// blk covers checkerboard sites blk*256 ..; a (z, t) plane holds P of them
// (P = Lx Ly / 512).  Die x owns the piece (y part x % py, z range x / py)
// of every plane and walks it plane by plane in t: all dies advance through
// t together, and a die's gathers from t +- 3 land in the 7 t-slices of its
// own piece it has just touched, which fit its L2 (the contiguous ranges of
// xcd_block put whole t-slices on a die: 7 of them do not fit).  A
// permutation of the logical blocks, so partial sums indexed by logical
// block do not change.
struct SlabOrder {
  int P, py, bpy, zs, Lz;   // blocks per plane, y pieces, blocks per piece per plane, z per piece
  int ny, nz, Lt;           // a die's piece swept as ny x nz sub-pieces, one t sweep each; t extent
};
__device__ __forceinline__ int slab_block(int b, const SlabOrder &s) {
  const int x = b % NXCD, j = b / NXCD;
  const int hy = x % s.py, zq = x / s.py;
  const int bs = s.bpy / s.ny, zsub = s.zs / s.nz;   // a sub-piece: bs blocks per plane, zsub planes
  const int per_t = zsub * bs, per_sweep = s.Lt * per_t;
  const int sub = j / per_sweep, jj = j - sub * per_sweep;
  const int sy = sub % s.ny, sz = sub / s.ny;
  const int t = jj / per_t, r = jj - t * per_t;
  const int zl = sz * zsub + r / bs, pb = sy * bs + (r - (r / bs) * bs);
  return (t * s.Lz + zq * s.zs + zl) * s.P + hy * s.bpy + pb;
}
// the slab order for g, or false when the lattice does not split that way
// (planes not a whole number of blocks, or no split of the 8 dies into y and
// z pieces that divides both); blocks of run sites of one parity; one sweep
// per die (ny = nz = 1)
bool slab_order(const Geom &g, SlabOrder &s, int run = 256);
// sweep each die's piece as ny x nz sub-pieces (y blocks, z planes), each
// through all t before the next: a smaller t-window per die; false (s
// unchanged) unless both divide the piece
bool slab_split(SlabOrder &s, int ny, int nz);

// --------------------------------------------------------------- precisions
// Complex loads/stores for the three storage precisions.  Accumulation is in
// double for FP64 storage and float for FP32/FP16 storage.
template <int P> struct Prec;
template <> struct Prec<QAMD_F64> {
  typedef double2 S;   // stored complex
  typedef double R;    // arithmetic type
  typedef double2 C;
};
template <> struct Prec<QAMD_F32> {
  typedef float2 S;
  typedef float R;
  typedef float2 C;
};
template <> struct Prec<QAMD_F16> {
  typedef __half2 S;
  typedef float R;
  typedef float2 C;
};
// Q16: the stored integers are returned unscaled by to_c; the 1/32767 and
// the norm / link scale are applied by the loaders
template <> struct Prec<QAMD_Q16> {
  typedef short2 S;
  typedef float R;
  typedef float2 C;
};
constexpr float Q16_ONE = 32767.0f;
__device__ __forceinline__ bool is_16bit(int p) { return p == QAMD_F16 || p == QAMD_Q16; }

__device__ __forceinline__ double2 to_c(double2 v, double) { return v; }
__device__ __forceinline__ float2 to_c(float2 v, float) { return v; }
__device__ __forceinline__ float2 to_c(__half2 v, float) { return __half22float2(v); }
__device__ __forceinline__ float2 to_c(short2 v, float) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ short2 q16_pack(float a, float b) {   // a, b in [-1, 1]
  return make_short2((short)__float2int_rn(a * Q16_ONE), (short)__float2int_rn(b * Q16_ONE));
}

typedef double dvec2_t __attribute__((ext_vector_type(2)));
typedef float fvec2_t __attribute__((ext_vector_type(2)));
// one store of a 16-, 8- or 4-byte value, non-temporal when NTS (an output
// whose lines would otherwise take cache space from data the kernel is still
// to read)
template <bool NTS, class T> __device__ __forceinline__ void put(T *p, const T &v) {
  if constexpr (!NTS) {
    *p = v;
  } else if constexpr (sizeof(T) == 16) {
    __builtin_nontemporal_store(*reinterpret_cast<const dvec2_t *>(&v), reinterpret_cast<dvec2_t *>(p));
  } else if constexpr (sizeof(T) == 8) {
    __builtin_nontemporal_store(*reinterpret_cast<const fvec2_t *>(&v), reinterpret_cast<fvec2_t *>(p));
  } else {
    static_assert(sizeof(T) == 4, "16-, 8- or 4-byte stores");
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned *>(&v), reinterpret_cast<unsigned *>(p));
  }
}

// non-temporal (streaming) loads of data read exactly once
__device__ __forceinline__ double2 ntload(const double2 *p) {
  dvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const dvec2_t *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ntload(const float2 *p) {
  fvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const fvec2_t *>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ short2 ntload(const short2 *p) {
  unsigned u = __builtin_nontemporal_load(reinterpret_cast<const unsigned *>(p));
  return make_short2((short)(u & 0xffff), (short)(u >> 16));
}
__device__ __forceinline__ __half2 ntload(const __half2 *p) {
  unsigned u = __builtin_nontemporal_load(reinterpret_cast<const unsigned *>(p));
  return *reinterpret_cast<__half2 *>(&u);
}

// one 16-byte non-temporal load: element k of the G terms of a link group
typedef float fvec4_t __attribute__((ext_vector_type(4)));
typedef unsigned uvec4_t __attribute__((ext_vector_type(4)));
// the G terms of one link group from a 16-byte word (same byte order as the
// grouped link layout)
template <class S>
__device__ __forceinline__ void unpack_group(const uvec4_t &v, S (&o)[LinkG<S>::G]) {
  static_assert(sizeof(o) == 16, "a link group is 16 bytes");
  __builtin_memcpy(o, &v, 16);
}
__device__ __forceinline__ void ntload_group(const double2 *p, double2 (&o)[1]) { o[0] = ntload(p); }
__device__ __forceinline__ void ntload_group(const float2 *p, float2 (&o)[2]) {
  fvec4_t v = __builtin_nontemporal_load(reinterpret_cast<const fvec4_t *>(p));
  o[0] = make_float2(v.x, v.y);
  o[1] = make_float2(v.z, v.w);
}
__device__ __forceinline__ void ntload_group(const short2 *p, short2 (&o)[4]) {
  uvec4_t v = __builtin_nontemporal_load(reinterpret_cast<const uvec4_t *>(p));
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = make_short2((short)(v[i] & 0xffff), (short)(v[i] >> 16));
}
__device__ __forceinline__ void ntload_group(const __half2 *p, __half2 (&o)[4]) {
  uvec4_t v = __builtin_nontemporal_load(reinterpret_cast<const uvec4_t *>(p));
#pragma unroll
  for (int i = 0; i < 4; i++) {
    unsigned u = v[i];
    o[i] = *reinterpret_cast<__half2 *>(&u);
  }
}
// the same 16 bytes through the ordinary (cached) load path
template <class S>
__device__ __forceinline__ void cload_group(const S *p, S (&o)[LinkG<S>::G]) {
  const uvec4_t v = *reinterpret_cast<const uvec4_t *>(p);
  unpack_group(v, o);
}
// A link group through an allocating load (ALLOC) or a non-temporal one.
// ALLOC must be a compile-time constant: two loads of one address that
// differ only in the non-temporal hint are merged into one without it.
template <bool ALLOC, class S>
__device__ __forceinline__ void link_group_load(const S *p, S (&o)[LinkG<S>::G]) {
  if constexpr (ALLOC) cload_group(p, o);
  else ntload_group(p, o);
}

// While one of these lives (a solver whose vectors fill the 256 MB MALL, such
// as multishift CG with two vectors per shift), the staggered Dslash loads
// every link non-temporally instead of keeping a share of them in the MALL.
struct MallShareOff {
  MallShareOff();
  ~MallShareOff();
  MallShareOff(const MallShareOff &) = delete;
  MallShareOff &operator=(const MallShareOff &) = delete;
};

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>): a loop
// whose index is a compile-time constant inside the body
template <class F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------- fields
struct Field {             // one parity of a colour-vector / spinor field
  int prec;                // QAMD_F64/F32/F16
  int ncomp;               // complex components per site and rhs (3 or 12)
  int nrhs;
  int parity;              // parity the field lives on (0/1)
  int rhs_inner;           // 1: spinor layout, 16-byte chunks of kc components, rhs next
  int kc;                  // components per 16-byte chunk: 1 (F64), 2 (F32), 4 (F16/Q16)
  int64_t vp;              // sites per parity, padded to 64
  const Geom *g;           // device copy of the geometry
  void *data;              // complex numbers, see cidx()
  float *norm;             // FP16 only, see nidx()
  __host__ __device__ int64_t elems() const { return (int64_t)nrhs * ncomp * vp; }
  // complex component j of rhs r at checkerboard site cb.  Spinors
  // (rhs_inner): [j / kc][cb][r][j % kc], so that one thread's components
  // come in 16-byte loads and the threads of one site (consecutive r) read
  // consecutive chunks; colour vectors: [r][j][cb].
  __host__ __device__ int64_t cidx(int r, int j, int64_t cb) const {
    return rhs_inner ? ((((int64_t)(j / kc) * vp + cb) * nrhs + r) * kc + j % kc)
                     : ((int64_t)r * ncomp + j) * vp + cb;
  }
  // 16-byte chunk q (components q*kc ... q*kc+kc-1) of rhs r at site cb
  __host__ __device__ int64_t chunk(int r, int q, int64_t cb) const { return ((int64_t)q * vp + cb) * nrhs + r; }
  __host__ __device__ int64_t nidx(int r, int64_t cb) const {
    return rhs_inner ? cb * nrhs + r : (int64_t)r * vp + cb;
  }
};

// -------------------------------------------------------------- reductions
// Deterministic grid reductions in two launches: every block of the reducing
// kernel stores its block sums (plain stores, no fences or atomics), then a
// one-block finisher kernel adds them in block order and runs the scalar
// epilogue.  The kernel boundary makes the partial sums visible, which on a
// multi-die GPU is far cheaper than a device-scope fence per block.
constexpr int MAXRED = 32;   // values reduced by one kernel
constexpr int NSETS = 4;     // independent partial-sum buffers
struct RedScratch {
  double *part;              // [MAXRED][max_blocks] of one set
  int max_blocks;
};
RedScratch red_set(int set);   // host: scratch set 0..NSETS-1 (reserve first)
int red_reserve(int nblocks);

template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int i = 0; i < N; i++) v[i] += __shfl_xor(v[i], o, 64);
}

// Reduce v over the block; thread 0 gets the block total.  blockDim.x must be
// a multiple of 64 and at most 1024.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
  __shared__ double sh[16][N];
  wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int nw = blockDim.x >> 6;
  if (l == 0)
#pragma unroll
    for (int i = 0; i < N; i++) sh[w][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      double s = 0;
      for (int k = 0; k < nw; k++) s += sh[k][i];
      v[i] = s;
    }
  }
  __syncthreads();
}

// block sums of v into slot idx of the partial-sum set
template <int N>
__device__ __forceinline__ void grid_partial_at(double (&v)[N], RedScratch rs, int idx) {
  block_sum(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < N; i++) rs.part[i * rs.max_blocks + idx] = v[i];
}
// block sums of v into the partial-sum set; every block calls it once
template <int N>
__device__ __forceinline__ void grid_partial(double (&v)[N], RedScratch rs) {
  grid_partial_at(v, rs, blockIdx.x);
}

// fixed-order total of N slots over nb blocks: thread k adds blocks k,
// k+FT, k+2FT, ... in order (all of a chunk's loads issued before its adds,
// so the round trips overlap), then the block sum adds the threads in order
constexpr int FT = 512;   // finisher threads
template <int N>
__device__ __forceinline__ void finish_sum(RedScratch rs, int nb, double (&w)[N]) {
  constexpr int U = N >= 16 ? 1 : 16 / N;
#pragma unroll
  for (int i = 0; i < N; i++) w[i] = 0.0;
  for (int b0 = threadIdx.x; b0 < nb; b0 += U * FT) {
    double v[U][N];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int i = 0; i < N; i++) {
        int b = b0 + u * FT;
        v[u][i] = b < nb ? rs.part[i * rs.max_blocks + b] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int i = 0; i < N; i++) w[i] += v[u][i];
  }
  block_sum(w);
}

template <int N, class Fin>
__global__ void __launch_bounds__(FT) k_finish(RedScratch rs, int nb, Fin fin, const int *skip) {
  if (skip && *skip) return;
  double w[N];
  finish_sum(rs, nb, w);
  if (threadIdx.x == 0) fin(w);
}

// totals of N0 slots of set a and N1 slots of set b, passed as one array
template <int N0, int N1, class Fin>
__global__ void __launch_bounds__(FT) k_finish2(RedScratch a, RedScratch b, int nb, Fin fin, const int *skip) {
  if (skip && *skip) return;
  double w0[N0], w1[N1], t[N0 + N1];
  finish_sum(a, nb, w0);
  finish_sum(b, nb, w1);
  if (threadIdx.x == 0) {
    for (int i = 0; i < N0; i++) t[i] = w0[i];
    for (int i = 0; i < N1; i++) t[N0 + i] = w1[i];
    fin(t);
  }
}

template <int N, class Fin>
inline void launch_finish(RedScratch rs, int nb, Fin fin, hipStream_t st, const int *skip = nullptr) {
  hipLaunchKernelGGL((k_finish<N, Fin>), dim3(1), dim3(FT), 0, st, rs, nb, fin, skip);
}

}  // namespace qamd
