// Wilson even-odd Dslash for gfx950 (lib/wilsonDslash.c:21-179, 272-285):
//   out = a*w + b * sum_t (1 + s_t g_mu) U_t(x) psi(x + d_t)
// term 2mu: U_mu(x), psi(x+mu), s = -1;  term 2mu+1: U_mu(x-mu)^+, psi(x-mu),
// s = +1 (spinproj{mu}m / spinproj{mu}p of lib/vlaw0.c.lua:545-649).  QLL's
// operator is a = massFac*(4+m) with w = y, b = -1.
//
// DeGrand-Rossi basis (lib/Gamma.lua:7-11): g_mu = [[0, A_mu], [A_mu^+, 0]],
// so (1 + s g) psi has upper half h = psi_up + s A psi_dn and lower half
// s A^+ h.  Each term therefore multiplies only the 2-spin half spinor h by
// the link (2 mat-vecs instead of 4) and reconstructs the lower half.
//
// One thread per (site, rhs); spinors are stored in 16-byte chunks with the
// rhs next to each other, so with several right-hand sides the lanes of a
// wavefront that share a site load the same link elements (each link is
// fetched once for all of them) and read consecutive chunks.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"

#include <stdio.h>
#include <string.h>

namespace qamd {

template <int P> struct WArgs {
  typedef typename Prec<P>::S S;
  Geom g;
  Field in, out, w;
  int has_w;
  const S *links;
  double a, b;
  int par;
  float lscale[8];
  RedScratch rs;
  double *red_out;
  const int *skip;
  int dag;     // 1: the hermitian conjugate g5 H g5 (every projector sign flipped)
  int xcd;     // block order: 1 XCD-aware (each die takes one contiguous range of sites), 2 slab sweep
  int tiled;   // sites in the tiled order of tl (else checkerboard order)
  Tiling tl;
  SlabOrder slab;   // xcd 2: the runs of 256 / nrhs sites in slab-sweep order
  const uint8_t *lsign;   // set: third rows rebuilt (determinant sign bit per site and term)
  Field rv;               // EPI 2: the residual updated in place, r -= alpha out
  const double *alpha;    // EPI 2: per-rhs step (device)
};

template <class C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }

// One 16-byte chunk of a spinor: K complex components of one (site, rhs)
// (K = 1 FP64, 2 FP32, 4 FP16/Q16), scaled by the site norm n for 16-bit
// storage.  A wavefront's vector-memory instructions are issued at a fixed
// rate whatever their width, so full 16-byte lanes carry the most data.
// Stores are non-temporal: the output no longer takes L2 lines from the
// neighbour spinors the kernel gathers (FP64 1-rhs Dslash 4 %, FP32 16-rhs
// 3 % faster at 32^3x64).
template <int P> struct Chunk;
template <> struct Chunk<QAMD_F64> {
  static constexpr int K = 1;
  __device__ static void ld(const void *d, int64_t i, float, double2 (&v)[1]) { v[0] = ((const double2 *)d)[i]; }
  __device__ static void st(void *d, int64_t i, const double2 (&v)[1]) { put<true>((double2 *)d + i, v[0]); }
};
template <> struct Chunk<QAMD_F32> {
  static constexpr int K = 2;
  __device__ static void ld(const void *d, int64_t i, float, float2 (&v)[2]) {
    fvec4_t u = ((const fvec4_t *)d)[i];
    v[0] = make_float2(u.x, u.y);
    v[1] = make_float2(u.z, u.w);
  }
  __device__ static void st(void *d, int64_t i, const float2 (&v)[2]) {
    fvec4_t u = {v[0].x, v[0].y, v[1].x, v[1].y};
    __builtin_nontemporal_store(u, (fvec4_t *)d + i);
  }
};
template <> struct Chunk<QAMD_F16> {
  static constexpr int K = 4;
  __device__ static void ld(const void *d, int64_t i, float n, float2 (&v)[4]) {
    uvec4_t u = ((const uvec4_t *)d)[i];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      unsigned w = u[k];
      float2 h = __half22float2(*reinterpret_cast<__half2 *>(&w));
      v[k] = make_float2(h.x * n, h.y * n);
    }
  }
  __device__ static void st(void *d, int64_t i, const float2 (&v)[4]) {
    uvec4_t u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      __half2 h = __floats2half2_rn(v[k].x, v[k].y);
      u[k] = *reinterpret_cast<unsigned *>(&h);
    }
    __builtin_nontemporal_store(u, (uvec4_t *)d + i);
  }
};
template <> struct Chunk<QAMD_Q16> {
  static constexpr int K = 4;
  __device__ static void ld(const void *d, int64_t i, float n, float2 (&v)[4]) {
    uvec4_t u = ((const uvec4_t *)d)[i];
    n *= 1.0f / Q16_ONE;
#pragma unroll
    for (int k = 0; k < 4; k++)
      v[k] = make_float2((float)(short)(u[k] & 0xffff) * n, (float)(short)(u[k] >> 16) * n);
  }
  __device__ static void st(void *d, int64_t i, const float2 (&v)[4]) {
    uvec4_t u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      short2 q = q16_pack(v[k].x, v[k].y);
      u[k] = (unsigned)(unsigned short)q.x | ((unsigned)(unsigned short)q.y << 16);
    }
    __builtin_nontemporal_store(u, (uvec4_t *)d + i);
  }
};

struct FinStore1 {
  double *dst;
  __device__ void operator()(const double *t) const { dst[0] = t[0]; }
};

// most sites one 256-thread block covers when links are staged in LDS
// (8 or more rhs per site)
constexpr int WMAXS = 256 / 8 + 1;

// REC: the links' third rows are not read but rebuilt from the first two
// (every term is a scale times a unitary matrix; 8·96 instead of 8·144 link
// bytes per site in FP64).  Used for one or two rhs, LDS staging for eight or
// more: never both.  EPI (with RED): 0 one sum |out|^2 over the block; 1 per-
// rhs sums |out_r|^2 (slot r); 2 the solver step r_r -= alpha_r out_r with out
// not stored, per-rhs sums |r_r|^2.
template <int P, bool RED, bool LDSL, bool REC, int EPI = 0>
__global__ void __launch_bounds__(256) k_wilson(WArgs<P> A) {
  static_assert(!(LDSL && REC), "the LDS link stage reads whole links");
  static_assert(EPI == 0 || RED, "per-rhs epilogues reduce");
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::S S;
  if (A.skip && *A.skip) return;
  // local copies: the kernel reads its argument structs only here
  const Field fin = A.in, fw = A.w;
  const Field fout = A.out;
  const Geom g = A.g;
  const int nr = fin.nrhs;
  const int blk = A.xcd == 2 ? slab_block(blockIdx.x, A.slab) : A.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  int64_t gid = (int64_t)blk * blockDim.x + threadIdx.x;
  const int li = (int)(gid / nr), r = (int)(gid - (int64_t)li * nr);
  constexpr int G = LinkG<S>::G;
  constexpr int NCH = 72 / G;   // 16-byte link words per site (8 terms x 9 elements)
  [[maybe_unused]] const uvec4_t *ls = nullptr;
  if constexpr (LDSL) {
    // links of the block's sites, loaded once into LDS: the nr threads of a
    // site then read them there instead of each issuing the same loads
    __shared__ uvec4_t lds[WMAXS * NCH];
    const int li0 = (int)(((int64_t)blk * blockDim.x) / nr);
    int li1 = (int)(((int64_t)blk * blockDim.x + blockDim.x - 1) / nr);
    li1 = li1 < g.vh - 1 ? li1 : (int)g.vh - 1;
    const int ns = li1 - li0 + 1;
    for (int k = threadIdx.x; k < ns * NCH; k += blockDim.x) {
      const int sl = k % ns, ch = k / ns;   // neighbouring threads: neighbouring sites
      const Site ss = A.tiled ? site_of_tiled(g, A.tl, li0 + sl, A.par) : site_of(g, li0 + sl, A.par);
      const S *lps = link_base<S, 8>(A.links, ss.lex >> 1) + link_off<S, 8>((ch / 9) * G, ch % 9);
      lds[sl * NCH + ch] = __builtin_nontemporal_load(reinterpret_cast<const uvec4_t *>(lps));
    }
    __syncthreads();
    ls = lds + (li - li0) * NCH;
  }
  double red[1] = {0.0};
  if (li < g.vh) {
    const Site s = A.tiled ? site_of_tiled(g, A.tl, li, A.par) : site_of(g, li, A.par);
    const int cb = s.lex >> 1;
    constexpr int K = Chunk<P>::K;
    const S *lp = link_base<S, 8>(A.links, cb);
    [[maybe_unused]] unsigned lsb = 0;
    if constexpr (REC) lsb = A.lsign[cb];
    C acc[4][3];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[a][c].x = acc[a][c].y = (R)0;
    // one link group: G terms sharing one 16-byte load per link element
    auto group = [&](const int tg) __attribute__((always_inline)) {
    S q[9][G];
    if constexpr (LDSL) {
#pragma unroll
      for (int e = 0; e < 9; e++) unpack_group(ls[tg * 9 + e], q[e]);
    } else {
#pragma unroll
      for (int e = 0; e < 9; e++)
        if (!(REC && e % 3 == 2)) ntload_group(lp + link_off<S, 8>(tg * G, e), q[e]);
    }
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
      const int t = tg * G + gg;
      const int mu = t >> 1;
      const bool bwd = t & 1;
      int nb = nbr_cb_rt(g, s, mu, bwd ? -1 : 1);
      // projector sign s = +1 for backward terms (-1 forward); the conjugate
      // flips it, g5 (1 -+ g_mu) g5 = (1 +- g_mu).  Multiplying by +-1 and
      // adding in one fma is exact, the same as adding or subtracting.
      const R sg = (bwd != (A.dag != 0)) ? (R)1 : (R)-1;
      float nrm = 1.0f;
      if constexpr (P == QAMD_F16 || P == QAMD_Q16) nrm = fin.norm[fin.nidx(r, nb)] * A.lscale[t];
      C psi[4][3];
#pragma unroll
      for (int q = 0; q < 12 / K; q++) {
        C v[K];
        Chunk<P>::ld(fin.data, fin.chunk(r, q, nb), nrm, v);
#pragma unroll
        for (int k = 0; k < K; k++) psi[(q * K + k) / 3][(q * K + k) % 3] = v[k];
      }
      // h = psi_up + s A psi_dn, s = +1 for backward terms.  A psi_dn per mu:
      //   mu 0: (i p3, i p2)  mu 1: (-p3, p2)  mu 2: (i p2, -i p3)  mu 3: (p2, p3)
      // written as component selects (no branches over mu)
      C h[2][3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const C p2 = psi[2][c], p3 = psi[3][c];
        C d0, d1;
        d0.x = sel4(mu, -p3.y, -p3.x, -p2.y, p2.x);
        d0.y = sel4(mu, p3.x, -p3.y, p2.x, p2.y);
        d1.x = sel4(mu, -p2.y, p2.x, p3.y, p3.x);
        d1.y = sel4(mu, p2.x, p2.y, -p3.x, p3.y);
        h[0][c].x = fma(sg, d0.x, psi[0][c].x);
        h[0][c].y = fma(sg, d0.y, psi[0][c].y);
        h[1][c].x = fma(sg, d1.x, psi[1][c].x);
        h[1][c].y = fma(sg, d1.y, psi[1][c].y);
      }
      C m[9];
#pragma unroll
      for (int e = 0; e < 9; e++)
        if (!(REC && e % 3 == 2)) m[e] = to_c(q[e][gg], (R)0);
      if constexpr (REC) long_row3<C, R>(m, (lsb >> t) & 1);
      C uh[2][3];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int rr = 0; rr < 3; rr++) {
          C z;
          z.x = z.y = (R)0;
#pragma unroll
          for (int c = 0; c < 3; c++) {
            z.x = fma(m[3 * c + rr].x, h[a][c].x, z.x);
            z.x = fma(-m[3 * c + rr].y, h[a][c].y, z.x);
            z.y = fma(m[3 * c + rr].x, h[a][c].y, z.y);
            z.y = fma(m[3 * c + rr].y, h[a][c].x, z.y);
          }
          uh[a][rr] = z;
        }
      // upper += Uh; lower += s A^+ Uh
#pragma unroll
      for (int c = 0; c < 3; c++) {
        acc[0][c] = cadd(acc[0][c], uh[0][c]);
        acc[1][c] = cadd(acc[1][c], uh[1][c]);
        // A^+ chi per mu:  mu 0: (-i u1, -i u0)  mu 1: (u1, -u0)
        //                  mu 2: (-i u0, i u1)   mu 3: (u0, u1)
        const C u0 = uh[0][c], u1 = uh[1][c];
        C l0, l1;
        l0.x = sel4(mu, u1.y, u1.x, u0.y, u0.x);
        l0.y = sel4(mu, -u1.x, u1.y, -u0.x, u0.y);
        l1.x = sel4(mu, u0.y, -u0.x, -u1.y, u1.x);
        l1.y = sel4(mu, -u0.x, -u0.y, u1.x, u1.y);
        acc[2][c].x = fma(sg, l0.x, acc[2][c].x);
        acc[2][c].y = fma(sg, l0.y, acc[2][c].y);
        acc[3][c].x = fma(sg, l1.x, acc[3][c].x);
        acc[3][c].y = fma(sg, l1.y, acc[3][c].y);
      }
    }
    };
    // FP64/FP32 (G <= 2): the groups are unrolled, so every term's direction
    // is a compile-time constant, with a scheduling barrier between groups
    // (without it the compiler hoists every group's loads to the top and
    // spills); 16-bit (G = 4): the two groups stay a rolled loop, which keeps
    // the four terms' spinors of one group in registers
    if constexpr (G <= 2) {
#pragma unroll
      for (int tg = 0; tg < 8 / G; tg++) {
        __builtin_amdgcn_sched_barrier(0);
        group(tg);
      }
    } else {
#pragma unroll 1
      for (int tg = 0; tg < 8 / G; tg++) group(tg);
    }
    const R b = (R)A.b, a = (R)A.a;
    C o[4][3];
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        o[sp][c].x = b * acc[sp][c].x;
        o[sp][c].y = b * acc[sp][c].y;
      }
    if (A.has_w) {
      float wn = 1.0f;
      if constexpr (P == QAMD_F16 || P == QAMD_Q16) wn = fw.norm[fw.nidx(r, cb)];
#pragma unroll
      for (int q = 0; q < 12 / K; q++) {
        C wv[K];
        Chunk<P>::ld(fw.data, fw.chunk(r, q, cb), wn, wv);
#pragma unroll
        for (int k = 0; k < K; k++) {
          C &v = o[(q * K + k) / 3][(q * K + k) % 3];
          v.x = fma(a, wv[k].x, v.x);
          v.y = fma(a, wv[k].y, v.y);
        }
      }
    }
    constexpr bool NORMED = (P == QAMD_F16 || P == QAMD_Q16);
    if constexpr (EPI == 2) {   // o <- r - alpha o, stored to r
      const Field &fr = A.rv;
      const R al = (R)A.alpha[r];
      float rn = 1.0f;
      if constexpr (NORMED) rn = fr.norm[fr.nidx(r, cb)];
#pragma unroll
      for (int q = 0; q < 12 / K; q++) {
        C rvv[K];
        Chunk<P>::ld(fr.data, fr.chunk(r, q, cb), rn, rvv);
#pragma unroll
        for (int k = 0; k < K; k++) {
          C &v = o[(q * K + k) / 3][(q * K + k) % 3];
          v.x = fma(-al, v.x, rvv[k].x);
          v.y = fma(-al, v.y, rvv[k].y);
        }
      }
    }
    const Field &fst = (EPI == 2) ? A.rv : fout;
    float on = 1.0f;
    if (NORMED) {
      float mx = 0.f;
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int c = 0; c < 3; c++) mx = fmaxf(mx, fmaxf(fabsf((float)o[sp][c].x), fabsf((float)o[sp][c].y)));
      put<true>(fst.norm + fst.nidx(r, cb), mx);
      on = mx > 0.f ? 1.0f / mx : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 12 / K; q++) {
      C v[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        v[k] = o[(q * K + k) / 3][(q * K + k) % 3];
        if (RED) red[0] += (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
        if (NORMED) {
          v[k].x *= on;
          v[k].y *= on;
        }
      }
      Chunk<P>::st(fst.data, fst.chunk(r, q, cb), v);
    }
  }
  if constexpr (EPI > 0) {
    // per-rhs sums in a fixed order: thread q < nr adds the values of the
    // block's threads that hold rhs q, in thread order
    __shared__ double shr[256];
    shr[threadIdx.x] = red[0];
    __syncthreads();
    if ((int)threadIdx.x < nr) {
      const int q = threadIdx.x, r0 = (int)(((int64_t)blk * blockDim.x) % nr);   // rhs of thread 0
      double t = 0.0;
      for (int tl = (q - r0 + nr) % nr; tl < (int)blockDim.x; tl += nr) t += shr[tl];
      A.rs.part[q * A.rs.max_blocks + blk] = t;
    }
  } else if (RED) {
    grid_partial_at(red, A.rs, blk);   // indexed by site block: same sums in either order
  }
}

// launch shapes the tuner tries: block order, links staged in LDS (8 or more
// rhs) and site tiling (log2 extents)
struct WShape {
  int xcd, lds, tile[4];
};
constexpr WShape WSHAPES[] = {
    {0, 0, {0, 0, 0, 0}}, {1, 0, {0, 0, 0, 0}}, {1, 0, {3, 3, 3, 3}}, {1, 0, {2, 2, 2, 2}},
    {1, 0, {4, 3, 3, 3}}, {1, 0, {3, 3, 2, 2}}, {1, 0, {2, 2, 2, 4}}, {1, 0, {5, 2, 2, 2}},
    {0, 1, {0, 0, 0, 0}}, {1, 1, {0, 0, 0, 0}}, {1, 1, {3, 3, 2, 2}}, {1, 1, {5, 2, 2, 2}},
    {1, 1, {3, 3, 3, 3}}, {1, 1, {2, 2, 2, 2}}, {1, 1, {4, 2, 2, 2}}, {1, 1, {4, 3, 3, 2}},
    {2, 0, {0, 0, 0, 0}}, {2, 1, {0, 0, 0, 0}},
};
constexpr int NWSHAPE = sizeof(WSHAPES) / sizeof(WSHAPES[0]);
// site tiling (log2 extents 32 x 4 x 4 x 4) of every reducing launch on a lattice that admits it
constexpr int WRTILE[4] = {5, 2, 2, 2};
constexpr bool wrtile_listed() {
  for (const WShape &sh : WSHAPES)
    if (!sh.lds && sh.tile[0] == WRTILE[0] && sh.tile[1] == WRTILE[1] && sh.tile[2] == WRTILE[2] &&
        sh.tile[3] == WRTILE[3])
      return true;
  return false;
}
static_assert(wrtile_listed(), "reducing launches need a shape with the fixed tiling and no LDS stage");

template <int P> struct WRun {
  WArgs<P> A;
  int nb, epi;
  hipStream_t st;
  int ncand, shape[NWSHAPE];
  template <bool RED, bool REC, int EPI = 0>
  static void launch_k(int lds, const WRun &R, const WArgs<P> &B) {
    if constexpr (REC) hipLaunchKernelGGL((k_wilson<P, RED, false, true, EPI>), dim3(R.nb), dim3(256), 0, R.st, B);
    else if (lds) hipLaunchKernelGGL((k_wilson<P, RED, true, false, EPI>), dim3(R.nb), dim3(256), 0, R.st, B);
    else hipLaunchKernelGGL((k_wilson<P, RED, false, false, EPI>), dim3(R.nb), dim3(256), 0, R.st, B);
  }
  static int run(int c, void *ctx) {
    WRun &R = *(WRun *)ctx;
    WArgs<P> B = R.A;
    const WShape &sh = WSHAPES[R.shape[c]];
    B.xcd = sh.xcd;
    if (sh.xcd == 2 && !slab_order(B.g, B.slab, 256 / B.in.nrhs)) return -1;
    B.tiled = sh.tile[0] ? 1 : 0;
    if (B.tiled && !B.tl.init(B.g.L, sh.tile)) return -1;
    if (R.epi == 1 || R.epi == 2) {   // per-rhs sums into B.rs, finished by the caller
      if (R.epi == 1) {
        if (B.lsign) launch_k<true, true, 1>(sh.lds, R, B);
        else launch_k<true, false, 1>(sh.lds, R, B);
      } else {
        if (B.lsign) launch_k<true, true, 2>(sh.lds, R, B);
        else launch_k<true, false, 2>(sh.lds, R, B);
      }
    } else if (B.red_out) {
      if (B.lsign) launch_k<true, true>(sh.lds, R, B);
      else launch_k<true, false>(sh.lds, R, B);
      launch_finish<1>(B.rs, R.nb, FinStore1{B.red_out}, R.st, B.skip);
    } else {
      if (B.lsign) launch_k<false, true>(sh.lds, R, B);
      else launch_k<false, false>(sh.lds, R, B);
    }
    QAMD_HIP(hipGetLastError());
    return 0;
  }
};

template <int P>
static int wlaunch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                   double *red_out, hipStream_t st, const int *skip, bool dag, int epi = 0,
                   const Field *rv = nullptr, const double *alpha = nullptr, RedScratch rsx = RedScratch{}) {
  typedef typename Prec<P>::S S;
  WArgs<P> A;
  memset(&A, 0, sizeof A);
  A.g = l->lat->g;
  A.in = in;
  A.out = out;
  if (w) {
    A.w = *w;
    A.has_w = 1;
  }
  A.links = (const S *)l->data[out.parity];
  A.a = a;
  A.b = b;
  A.par = out.parity;
  A.skip = skip;
  A.dag = dag ? 1 : 0;
  // rebuilding costs every thread of a site the same arithmetic: it pays for
  // one or two rhs (link bytes dominate: FP64 1 rhs 0.160 -> 0.139 ms at
  // 32^4) and not for blocks (16 rhs: FP32 +3 %, Q16 +18 % time)
  A.lsign = (l->recon && in.nrhs <= 2) ? l->lsign[out.parity] : nullptr;
  for (int t = 0; t < 8; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  int64_t nthreads = A.g.vh * in.nrhs;
  WRun<P> R;
  R.A = A;
  R.nb = (int)((nthreads + 255) / 256);
  R.st = st;
  R.epi = epi;
  if (epi) {
    R.A.rs = rsx;
    if (epi == 2) {
      R.A.rv = *rv;
      R.A.alpha = alpha;
    }
  } else if (red_out) {
    if (red_reserve(R.nb)) return -1;
    R.A.rs = red_set(0);
    R.A.red_out = red_out;
  }
  // with a reduction every candidate groups the sites into blocks the same
  // way, so the partial sums (and their rounding) do not depend on the
  // tuner's pick: the fixed tiling WRTILE where the lattice admits it, else
  // checkerboard order.  LDS staging needs at least 8 rhs per site (at most
  // WMAXS sites per block)
  const bool rsum = red_out != nullptr || epi != 0;
  Tiling rt;
  const bool rtiled = rsum && rt.init(A.g.L, WRTILE);
  R.ncand = 0;
  for (int k = 0; k < NWSHAPE; k++) {
    const WShape &sh = WSHAPES[k];
    if (sh.lds && in.nrhs < 8) continue;
    if (sh.xcd == 2) {   // blocks are runs of whole sites, in planes of whole runs
      SlabOrder so;
      if (256 % in.nrhs != 0 || !slab_order(A.g, so, 256 / in.nrhs)) continue;
    }
    if (rsum) {
      bool fixed = true;
      for (int i = 0; i < 4; i++) fixed = fixed && sh.tile[i] == WRTILE[i];
      if (rtiled ? !fixed : sh.tile[0] != 0) continue;
    } else {
      Tiling t;
      if (sh.tile[0] && !t.init(A.g.L, sh.tile)) continue;
    }
    R.shape[R.ncand++] = k;
  }
  const int def = R.ncand > 1 ? 1 : 0;
  // the tuner's trial launches must not feed the output back into an input,
  // nor repeat an in-place residual update
  const bool alias = out.data == in.data || (w && out.data == w->data) || epi == 2;
  char key[96];
  snprintf(key, sizeof key, "wilson p%d n%d red%d %dx%dx%dx%d", P, in.nrhs, red_out || epi ? 1 : 0, A.g.L[0],
           A.g.L[1], A.g.L[2], A.g.L[3]);
  int c = alias ? def : tune_pick(key, R.ncand, def, &WRun<P>::run, &R);
  return WRun<P>::run(c, &R);
}

int wilson_dslash_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                         double *red_out, hipStream_t st, const int *skip, bool dag) {
  if (l->kind != 1) {
    set_error("wilson_dslash: links are not Wilson links");
    return -1;
  }
  if (out.prec != l->prec || in.prec != l->prec || (w && w->prec != l->prec)) {
    set_error("wilson_dslash: field precision does not match the links");
    return -1;
  }
  if (out.ncomp != 12 || in.ncomp != 12 || out.nrhs != in.nrhs || (w && w->nrhs != in.nrhs)) {
    set_error("wilson_dslash: needs spinor fields with equal rhs counts");
    return -1;
  }
  if (out.parity == in.parity || (w && w->parity != out.parity)) {
    set_error("wilson_dslash: parity mismatch");
    return -1;
  }
  switch (l->prec) {
    case QAMD_F64: return wlaunch<QAMD_F64>(l, out, in, a, w, b, red_out, st, skip, dag);
    case QAMD_F32: return wlaunch<QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, dag);
    case QAMD_F16: return wlaunch<QAMD_F16>(l, out, in, a, w, b, red_out, st, skip, dag);
    default: return wlaunch<QAMD_Q16>(l, out, in, a, w, b, red_out, st, skip, dag);
  }
}

// The Wilson solver's fused launches: out = a w + b H in with per-rhs sums
// |out_r|^2 (epi 1), or r_r -= alpha_r (a w + b H in)_r without storing out
// and per-rhs sums |r_r|^2 (epi 2); the sums go to slot r of rs over the
// launch's blocks, ceil(vh * nrhs / 256), for the caller to finish.
int wilson_dslash_rhs_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                             hipStream_t st, bool dag, int epi, Field *r, const double *alpha, RedScratch rs) {
  if (l->kind != 1 || in.prec != l->prec || (w && w->prec != l->prec) || in.ncomp != 12 ||
      (w && (w->nrhs != in.nrhs || w->parity == in.parity)) || (epi != 1 && epi != 2) ||
      (epi == 1 && (out.prec != l->prec || out.nrhs != in.nrhs || out.parity == in.parity)) ||
      (epi == 2 && (!r || !alpha || r->prec != l->prec || r->nrhs != in.nrhs || r->parity == in.parity)) ||
      in.nrhs > 16) {
    set_error("wilson_dslash_rhs: field shapes or precisions do not fit");
    return -1;
  }
  if (red_reserve((int)((l->lat->g.vh * in.nrhs + 255) / 256))) return -1;
  Field &o = epi == 2 ? *r : out;
  switch (l->prec) {
    case QAMD_F64: return wlaunch<QAMD_F64>(l, o, in, a, w, b, nullptr, st, nullptr, dag, epi, r, alpha, rs);
    case QAMD_F32: return wlaunch<QAMD_F32>(l, o, in, a, w, b, nullptr, st, nullptr, dag, epi, r, alpha, rs);
    case QAMD_F16: return wlaunch<QAMD_F16>(l, o, in, a, w, b, nullptr, st, nullptr, dag, epi, r, alpha, rs);
    default: return wlaunch<QAMD_Q16>(l, o, in, a, w, b, nullptr, st, nullptr, dag, epi, r, alpha, rs);
  }
}

}  // namespace qamd
