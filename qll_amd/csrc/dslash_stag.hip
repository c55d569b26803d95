// Asqtad / staggered even-odd Dslash for gfx950, and the link-set objects.
//
// One thread per output site of one parity.  The 16 (or 8) link matrices of
// a site are streamed from the wavefront-blocked link array with
// non-temporal loads (read once, kept out of L2), the neighbour colour
// vectors are gathered through L2/MALL, and the sum is accumulated in
// registers: out = a*w + b * sum_t L_t(x) in(x + d_t).
//
// Term order follows QLL's StaggeredLinks (lib/stagDslash.c:438-448, 820-826):
// with Naik, term 2k is the 1-hop link u[k] and 2k+1 the 3-hop link u3[k];
// k = 2mu gathers from x+mu, k = 2mu+1 from x-mu.  In storage the 1-hop terms
// come first (slots 0-7) and the 3-hop terms after them (slots 8-15).
//
// Forward-only link sets (QLL's createStaggeredLinksFb, lib/stagDslash.c:
// 464-519, applied by dslashFb2T, lib/stagDslash.c:110-176) store U_mu(x) in
// slot mu and L_mu(x) in slot 4 + mu for both parities, half the doublestored
// storage; the backward terms -U_mu(x-mu)^+ in(x-mu) and -L_mu(x-3mu)^+
// in(x-3mu) read the other parity's forward link at the neighbour, the same
// site index as the input vector they multiply.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"
#include "qamd_vec3.h"

#include <atomic>
#include <stdlib.h>
#include <string.h>
#include <math.h>

namespace qamd {

// links in precision P, input in precision PI, output and w in precision PO
// (PI = PO = P, or FP32 input and/or output with 16-bit links for the
// mixed-precision solvers)
template <int P, int PO = P, int PI = P> struct StagArgs {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PO>::S SO;
  typedef typename Prec<PI>::S SI;
  Geom g;
  const S *links;
  const SI *in;
  const float *in_norm;
  SO *out;
  float *out_norm;
  const SO *w;
  const float *w_norm;
  double a, b;
  int par;            // output parity
  float lscale[16];   // FP16 per-term link scale
  RedScratch rs;
  double *red_out;    // receives {|out|^2, |w|^2}
  const int *skip;    // optional device flag: whole grid returns when set
  const uint8_t *lsign;   // long-term determinant signs when their third rows are rebuilt
  const S *links_o;       // forward-only sets: the input parity's links (backward terms)
  const uint8_t *lsign_o; // and their long-term determinant signs
  // single-reduction CG step (CGX = 2, FP32 vectors, input t): Ar =
  // a r + b D t (not stored); p = r + beta p, s = Ar + beta s, x += alpha p,
  // r -= alpha s, |r|^2 -> red[0]; alpha and beta formed by every block from
  // the partial sums of |r|^2 (cg_rin) and |t|^2 (cg_tin), state in cg_state
  float2 *cgx, *cgr, *cgp, *cgs;
  CGScal *cg_state;
  int cg_ph, cg_nb;
  RedScratch cg_rin, cg_tin;
};

// The CG step's alpha and beta, formed by every block in the same fixed
// order (so all blocks agree bit for bit): thread k adds the partial sums of
// blocks k, k + blockDim, ... in order, then the block adds its threads in
// order; the recurrences are those of Chronopoulos & Gear (1989).
template <class Args>
__device__ __forceinline__ void cg_block_scalars(const Args &A, float &al_out, float &be_out, float &alp_out) {
  __shared__ double sab[3];
  double w[2] = {0.0, 0.0};
  for (int b0 = threadIdx.x; b0 < A.cg_nb; b0 += 4 * blockDim.x) {
    double v[4][2];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int b = b0 + u * blockDim.x;
      v[u][0] = b < A.cg_nb ? A.cg_rin.part[b] : 0.0;
      v[u][1] = b < A.cg_nb ? A.cg_tin.part[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      w[0] += v[u][0];
      w[1] += v[u][1];
    }
  }
  block_sum(w);
  if (threadIdx.x == 0) {
    CGScal *S = A.cg_state;
    const int ph = A.cg_ph;
    const double g = w[0], d = A.a * g + w[1];
    double be = 0.0, al = g / d;
    int keep_next = 0;
    double rho_next = 0.0, pap_next = 0.0;
    if (S->keep[ph] == 1) {
      // After a reliable update that keeps p: the true residual r is not
      // orthogonal to p_old, so the single-reduction identities do not hold
      // for this step and the next.  Here beta makes p = r + beta p_old
      // A-conjugate to p_old, p.Ap is formed from the measured products and
      // alpha is the exact line search (r_next orthogonal to p).
      const double kr = S->kr[ph], kp = S->kp[ph], kq = S->kq[ph];
      if (kp > 0.0) {
        const double b1 = -kr / kp, pap = d + b1 * kr;   // d + 2 b kr + b^2 kp with b = -kr/kp
        if (pap > 0.0) {
          be = b1;
          al = (g + b1 * kq) / pap;
          keep_next = 2;
          rho_next = g - al * (d + b1 * kr);   // r_next . r
          pap_next = pap;
        }
      }
      if (!keep_next) al = g / d;   // non-positive measured p.Ap: restart
    } else if (S->keep[ph] == 2) {
      // the step after: beta conjugate to the previous p (r.r_prev = rho is
      // not zero), alpha = gamma / p.Ap; from the next step on r is
      // orthogonal to the previous r and p, and the usual recurrences resume
      const double pp = S->pap[ph], b1 = (g - S->rho[ph]) / (S->al[ph] * pp), pap = d - b1 * b1 * pp;
      if (pap > 0.0) {
        be = b1;
        al = g / pap;
      }
    } else if (!S->restart[ph]) {
      be = g / S->gam[ph];
      al = g / (d - be * g / S->al[ph]);
    }
    sab[0] = al;
    sab[1] = be;
    sab[2] = S->al[ph];   // the previous iteration's alpha
    if (blockIdx.x == 0) {
      S->gam[1 - ph] = g;
      S->al[1 - ph] = al;
      S->restart[1 - ph] = 0;
      S->keep[1 - ph] = keep_next;
      S->rho[1 - ph] = rho_next;
      S->pap[1 - ph] = pap_next;
      S->rsq = g;
      S->tt = w[1];
      S->alpha = al;
      S->beta = be;
      S->maxrr = fmax(S->maxrr, sqrt(g));
      S->its += 1;
    }
  }
  __syncthreads();
  al_out = (float)sab[0];
  be_out = (float)sab[1];
  alp_out = (float)sab[2];
}

struct FinStore2 {
  double *dst;
  __device__ void operator()(const double *t) const {
    dst[0] = t[0];
    dst[1] = t[1];
  }
};

// input vector at site nb for the term in link slot t (16-bit links: the
// term's scale folded into the vector)
template <int P, int PO, int PI>
__device__ __forceinline__ void stag_vin(const StagArgs<P, PO, PI> &A, int64_t vp, int nb, int t,
                                         typename Prec<P>::C (&v)[3]) {
  typedef typename Prec<P>::R R;
  if constexpr (PI == P) {
    VecLd<P>::ld(A.in, A.in_norm, vp, nb, v, (P == QAMD_F16 || P == QAMD_Q16) ? A.lscale[t] : (R)1);
  } else {   // FP32 input, 16-bit links: the term's link scale applied here
    VecLd<PI>::ld(A.in, A.in_norm, vp, nb, v, (R)1);
    const R sc = A.lscale[t];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      v[j].x *= sc;
      v[j].y *= sc;
    }
  }
}

// acc -= m^+ v  (element (r,c) of m at m[3c+r])
template <class C>
__device__ __forceinline__ void cmac_adj_sub(C (&acc)[3], const C (&m)[9], const C (&v)[3]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const C a = m[3 * r + c];
      acc[r].x = fma(-a.x, v[c].x, acc[r].x);
      acc[r].x = fma(-a.y, v[c].y, acc[r].x);
      acc[r].y = fma(-a.x, v[c].y, acc[r].y);
      acc[r].y = fma(a.y, v[c].x, acc[r].y);
    }
}

// One slot of a forward-only set at the site whose stream starts at base
// (layout: link_off_fb): whole 16-byte words, the third-row words skipped
// when they are rebuilt (lng).  NT: non-temporal loads (a site's own forward
// links, read once); otherwise cached (a neighbour's, whose other slots
// other sites read).
template <class S, bool NT>
__device__ __forceinline__ uvec4_t fb_word(const S *p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const uvec4_t *>(p));
  else return *reinterpret_cast<const uvec4_t *>(p);
}
template <class S, bool NT>
__device__ __forceinline__ S fb_elem(const S *p) {
  if constexpr (NT) return ntload(p);
  else return *p;
}
template <int P, bool NT>
__device__ __forceinline__ void fb_load_slot(const typename Prec<P>::S *base, int sl, bool lng,
                                             typename Prec<P>::C (&m)[9]) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<P>::R R;
  constexpr int G = LinkG<S>::G;
  if constexpr (G == 1) {
#pragma unroll
    for (int e = 0; e < 9; e++)
      if (!(lng && e % 3 == 2)) m[e] = to_c(fb_elem<S, NT>(base + link_off_fb<S>(sl, e)), (R)0);
  } else if constexpr (G == 2) {   // words {0,1} {3,4} {6,7} {2,5}, then element 8
    S w[2];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      unpack_group(fb_word<S, NT>(base + link_off_fb<S>(sl, 3 * j)), w);
      m[3 * j] = to_c(w[0], (R)0);
      m[3 * j + 1] = to_c(w[1], (R)0);
    }
    if (!lng) {
      unpack_group(fb_word<S, NT>(base + link_off_fb<S>(sl, 2)), w);
      m[2] = to_c(w[0], (R)0);
      m[5] = to_c(w[1], (R)0);
      m[8] = to_c(fb_elem<S, NT>(base + link_off_fb<S>(sl, 8)), (R)0);
    }
  } else {   // G == 4: words {0,1,3,4} {6,7,2,5}, then element 8
    S w[4];
    unpack_group(fb_word<S, NT>(base + link_off_fb<S>(sl, 0)), w);
    m[0] = to_c(w[0], (R)0);
    m[1] = to_c(w[1], (R)0);
    m[3] = to_c(w[2], (R)0);
    m[4] = to_c(w[3], (R)0);
    const S *p1 = base + link_off_fb<S>(sl, 6);
    if (lng) {   // elements 6, 7: the leading 8 bytes of the second word
      uint64_t u;
      if constexpr (NT) u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(p1));
      else u = *reinterpret_cast<const uint64_t *>(p1);
      S h[2];
      __builtin_memcpy(h, &u, 8);
      m[6] = to_c(h[0], (R)0);
      m[7] = to_c(h[1], (R)0);
    } else {
      unpack_group(fb_word<S, NT>(p1), w);
      m[6] = to_c(w[0], (R)0);
      m[7] = to_c(w[1], (R)0);
      m[2] = to_c(w[2], (R)0);
      m[5] = to_c(w[3], (R)0);
      m[8] = to_c(fb_elem<S, NT>(base + link_off_fb<S>(sl, 8)), (R)0);
    }
  }
}

// Forward-only link set (8 storage slots: fat mu in slot mu, long mu in slot
// 4 + mu; FBT = 4 without long links, 8 with them):
//   acc = sum_mu U_mu(x) in(x+mu) - U_mu(x-mu)^+ in(x-mu)
//       + sum_mu L_mu(x) in(x+3mu) - L_mu(x-3mu)^+ in(x-3mu)
// The forward links are streamed from this site's block; a backward link is
// read at the neighbour (through the caches for FP32 and 16-bit sets, see
// below).  REC: long third rows rebuilt from the first two.
template <int P, int PO, int PI, int FBT, bool REC>
__device__ __forceinline__ void stag_acc_fb(const StagArgs<P, PO, PI> &A, const Site &s, int cb, int64_t vp,
                                            typename Prec<P>::C (&acc)[3]) {
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::S S;
  const S *lp = link_base<S, 8>(A.links, cb);
  [[maybe_unused]] unsigned lsb = 0;
  if constexpr (REC) lsb = A.lsign[cb];
#pragma unroll
  for (int sl = 0; sl < FBT; sl++) {
    const int mu = sl & 3, hop = sl < 4 ? 1 : 3;
    const bool lng = REC && sl >= 4;
    C m[9];
    fb_load_slot<P, true>(lp, sl, lng, m);
    C v[3];
    stag_vin(A, vp, nbr_cb(A.g, s, mu, hop), sl, v);
    if (lng) long_row3<C, R>(m, (lsb >> (sl - 4)) & 1);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) cmac<C, R>(acc[r], m[3 * c + r], v[c]);
  }
#pragma unroll
  for (int sl = 0; sl < FBT; sl++) {
    const int mu = sl & 3, hop = sl < 4 ? 1 : 3;
    const bool lng = REC && sl >= 4;
    const int nb = nbr_cb(A.g, s, mu, -hop);
    C m[9];
    // backward links through the caches for FP32 and 16-bit sets (the set
    // stays in the MALL from call to call: 32^4, non-temporal instead, FP32
    // 0.085 -> 0.100 ms, Q16 0.050 -> 0.059 ms), non-temporal for FP64 sets,
    // which do not fit it (0.188-0.195 -> 0.180-0.188 ms)
    fb_load_slot<P, P == QAMD_F64>(link_base<S, 8>(A.links_o, nb), sl, lng, m);
    C v[3];
    stag_vin(A, vp, nb, sl, v);
    if (lng) long_row3<C, R>(m, (A.lsign_o[nb] >> (sl - 4)) & 1);
    cmac_adj_sub(acc, m, v);
  }
}

// REC (16 terms): the long terms' third rows are not loaded but rebuilt
// from the first two (FP64: 2016 instead of 2400 bytes per site).
// FBT != 0: forward-only link set (NT = 8 storage slots), see stag_acc_fb.
template <int P, int PO, int NT, bool RED, bool REC = false, int CGX = 0, int PI = P, int FBT = 0, int RG = 0>
__device__ __forceinline__ void stag_dslash_body(const StagArgs<P, PO, PI> &A) {
  static_assert(PI == P || (PI == QAMD_F32 && (P == QAMD_F16 || P == QAMD_Q16)),
                "input precision: the links' or FP32 with 16-bit links");
  static_assert(!CGX || (PO == QAMD_F32 && RED), "fused CG step: FP32 vectors, with the |r|^2 sums");
  static_assert(!FBT || (NT == 8 && (FBT == 4 || FBT == 8) && (!REC || FBT == 8)), "forward-only link set");
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::S S;
  const int64_t vp = A.g.vh_pad;
  if (A.skip && *A.skip) return;   // uniform over the grid
  int blk = xcd_block(blockIdx.x, gridDim.x);
  int cb = blk * 256 + threadIdx.x;
  double red[2] = {0.0, 0.0};
  C acc[3];
#pragma unroll
  for (int j = 0; j < 3; j++) acc[j].x = acc[j].y = (R)0;
  if (cb < A.g.vh) {
    Site s = site_of(A.g, cb, A.par);
    constexpr int G = LinkG<S>::G;
    static_assert(!REC || NT == 16 || FBT, "long-row rebuild: 16-term or forward-only links");
    if constexpr (FBT) stag_acc_fb<P, PO, PI, FBT, REC>(A, s, cb, vp, acc);
    [[maybe_unused]] const S *lp = link_base<S, NT>(A.links, cb);
    [[maybe_unused]] const S *lp3 = link_base3<S, NT>(lp, vp);   // the 3-hop terms' region
    [[maybe_unused]] unsigned lsb = 0;
    if constexpr (REC && !FBT) lsb = A.lsign[cb];
    // slot group so (slots sg*G .. sg*G+G-1: all 1-hop terms or all 3-hop
    // ones; in a 16-term set the 3-hop groups are the odd positions of the
    // group order); the first RG 3-hop groups load through the caches
    auto group = [&](auto so_c) __attribute__((always_inline)) {
      constexpr int so = decltype(so_c)::value;
      constexpr int sg = group_at<G, NT>(so);
      constexpr bool H3 = NT == 16 && sg * G >= 8;
      constexpr bool ALLOC = H3 && so / 2 < RG;
      const bool lng = REC && H3;
      S q[9][G];   // element e of the group's G terms, one 16-byte load each
#pragma unroll
      for (int e = 0; e < 9; e++) {
        if (lng && e % 3 == 2) continue;   // third row of a long term: rebuilt below
        link_group_load<ALLOC>((H3 ? lp3 : lp) + link_off<S, NT>(sg * G, e), q[e]);
      }
#pragma unroll
      for (int gg = 0; gg < G; gg++) {
        const int sl = sg * G + gg;
        const int t = slot_term(sl, NT);
        const int k = (NT == 16) ? t / 2 : t;
        const int hop = (NT == 16 && (t & 1)) ? 3 : 1;
        const int mu = k / 2;
        const int h = (k & 1) ? -hop : hop;
        int nb = nbr_cb(A.g, s, mu, h);
        C v[3];
        stag_vin(A, vp, nb, t, v);
        C m[9];
#pragma unroll
        for (int e = 0; e < 9; e++)
          if (!(lng && e % 3 == 2)) m[e] = to_c(q[e][gg], (R)0);
        if (lng) long_row3<C, R>(m, (lsb >> (sl - 8)) & 1);
        // column major: element (r,c) = m[3c+r]
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) cmac<C, R>(acc[r], m[3 * c + r], v[c]);
      }
    };
    if constexpr (!FBT) static_for<NT / G>(group);
  }
  // the CG step's scalars: a block-wide sum, after the link stream
  [[maybe_unused]] float cg_al = 0.f, cg_be = 0.f, cg_alp = 0.f;
  if constexpr (CGX == 2) cg_block_scalars(A, cg_al, cg_be, cg_alp);
  if (cb < A.g.vh) {
    C o[3];
    const R b = (R)A.b;
    if constexpr (CGX == 2) {
      // x is updated every second iteration, for two: in iteration k + 1
      // (odd ph) x += alpha_k p_k + alpha_{k+1} p_{k+1}, with p_k the p read
      // here, so even iterations leave x alone (one vector stream fewer)
      float2 rv[3], pv[3], sv[3], xv[3], po[3];
      VecLd<QAMD_F32>::ld(A.cgr, nullptr, vp, cb, rv, 1.0f);
      VecLd<QAMD_F32>::ld(A.cgp, nullptr, vp, cb, pv, 1.0f);
      VecLd<QAMD_F32>::ld(A.cgs, nullptr, vp, cb, sv, 1.0f);
      const bool xup = A.cg_ph == 1;
      if (xup) VecLd<QAMD_F32>::ld(A.cgx, nullptr, vp, cb, xv, 1.0f);
      const float a = (float)A.a, al = cg_al, be = cg_be, alp = cg_alp;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const float wx = fmaf(a, rv[j].x, (float)b * acc[j].x), wy = fmaf(a, rv[j].y, (float)b * acc[j].y);
        po[j] = pv[j];
        pv[j].x = fmaf(be, pv[j].x, rv[j].x);
        pv[j].y = fmaf(be, pv[j].y, rv[j].y);
        sv[j].x = fmaf(be, sv[j].x, wx);
        sv[j].y = fmaf(be, sv[j].y, wy);
        rv[j].x = fmaf(-al, sv[j].x, rv[j].x);
        rv[j].y = fmaf(-al, sv[j].y, rv[j].y);
        red[0] += (double)rv[j].x * rv[j].x + (double)rv[j].y * rv[j].y;
      }
      VecSt<QAMD_F32>::st(A.cgp, nullptr, vp, cb, pv);
      VecSt<QAMD_F32>::st(A.cgs, nullptr, vp, cb, sv);
      VecSt<QAMD_F32>::st(A.cgr, nullptr, vp, cb, rv);
      if (xup) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          xv[j].x = fmaf(al, pv[j].x, fmaf(alp, po[j].x, xv[j].x));
          xv[j].y = fmaf(al, pv[j].y, fmaf(alp, po[j].y, xv[j].y));
        }
        VecSt<QAMD_F32>::st(A.cgx, nullptr, vp, cb, xv);
      }
    } else if (A.w) {
      C w[3];
      VecLd<PO>::ld(A.w, A.w_norm, vp, cb, w, (R)1);
      const R a = (R)A.a;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        o[j].x = fma(a, w[j].x, b * acc[j].x);
        o[j].y = fma(a, w[j].y, b * acc[j].y);
        if (RED) red[1] += (double)w[j].x * w[j].x + (double)w[j].y * w[j].y;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 3; j++) {
        o[j].x = b * acc[j].x;
        o[j].y = b * acc[j].y;
      }
    }
    if constexpr (!CGX) {
      if (RED) {
#pragma unroll
        for (int j = 0; j < 3; j++) red[0] += (double)o[j].x * o[j].x + (double)o[j].y * o[j].y;
      }
      VecSt<PO>::st(A.out, A.out_norm, vp, cb, o);
    }
  }
  if (RED) grid_partial<2>(red, A.rs);
}

template <int P, int PO, int NT, bool RED, bool REC = false, int CGX = 0, int PI = P, int FBT = 0, int RG = 0>
__global__ void __launch_bounds__(256) k_stag_dslash(StagArgs<P, PO, PI> A) {
  stag_dslash_body<P, PO, NT, RED, REC, CGX, PI, FBT, RG>(A);
}

// MI355X's 256 MB last-level cache (MALL) keeps lines loaded through the
// caches from one kernel launch to the next, while non-temporal loads stream
// past it.  16-term sets with rebuilt long rows load their first RG 3-hop
// slot groups through the caches -- at most 104 MB per FP32 or 16-bit set,
// 52 MB per FP64 set (whose solvers' vectors are twice as large) -- so that
// those bytes come from the MALL on the next call with the same links; a
// solver alternates the two parities' sets, and both sets' shares fit
// together with its vectors.  Measured at 32^4 (one 3-hop group = 50 MB;
// one set called repeatedly / the parities alternating / the solve):
//   FP64, share 0 -> 52 MB: Dslash 0.179 -> 0.176 ms / 0.179 -> 0.176 ms,
//     FP64 CG 0.759 -> 0.743 s (104 MB: 0.174 / 0.181 ms, CG 0.772 s);
//   FP32, 0 -> 104 MB: 0.089 -> 0.083-0.086 / 0.091 -> 0.086-0.088 ms,
//     mixed CG with FP32 links 0.388 -> 0.367-0.375 s (208 MB thrashes when
//     the parities alternate: 0.094 ms);
//   Q16, 0 -> 104 MB: 0.052 -> 0.047 ms either way, mixed CG with Q16 links
//     0.243 -> 0.222 s.
// Multishift CG keeps no share (MallShareOff): its 2 vectors per shift fill
// the MALL (12 shifts at 32^3x64, FP32 share 100 MB per set: 1.75 -> 1.79 s).
// qamd_set_mall_share, or QAMD_MALL_MB64 (FP64 sets) and QAMD_MALL_MB (FP32
// and 16-bit sets) in the environment, set the share per set in MB (0: every
// link load non-temporal).
static std::atomic<int> g_mall_off{0};
MallShareOff::MallShareOff() { g_mall_off++; }
MallShareOff::~MallShareOff() { g_mall_off--; }

static int g_mall_mb[2] = {-1, -1};   // FP64 sets, FP32/16-bit sets; -1: default
static int mall_mb(bool fp64) {
  int &mb = g_mall_mb[fp64 ? 0 : 1];
  if (mb < 0) {
    const char *e = getenv(fp64 ? "QAMD_MALL_MB64" : "QAMD_MALL_MB");
    mb = e ? atoi(e) : (fp64 ? 52 : 104);
    if (mb < 0) mb = 0;
  }
  return mb;
}
template <class S>
static int mall_groups(const Geom &g) {
  if (g_mall_off.load() > 0) return 0;
  const int64_t per = (int64_t)6 * 16 * g.vh;   // bytes of one 3-hop group (two rows of 16-byte words)
  int64_t n = (int64_t)mall_mb(LinkG<S>::G == 1) * 1000000 / per;
  const int ng = 8 / LinkG<S>::G;               // 3-hop groups of a 16-term set
  if (n > ng) n = ng;
  return n >= 4 ? 4 : n >= 2 ? 2 : (int)n;
}

template <int P, int PO, int NT, int PI, int FBT = 0>
static void launch_t(const StagArgs<P, PO, PI> &A, bool red, hipStream_t st) {
  int nb = (int)((A.g.vh + 255) / 256);
  if constexpr (FBT) {
    if (FBT == 8 && A.lsign) {
      if (red)
        hipLaunchKernelGGL((k_stag_dslash<P, PO, 8, true, FBT == 8, 0, PI, FBT>), dim3(nb), dim3(256), 0, st, A);
      else
        hipLaunchKernelGGL((k_stag_dslash<P, PO, 8, false, FBT == 8, 0, PI, FBT>), dim3(nb), dim3(256), 0, st, A);
      return;
    }
    if (red)
      hipLaunchKernelGGL((k_stag_dslash<P, PO, 8, true, false, 0, PI, FBT>), dim3(nb), dim3(256), 0, st, A);
    else
      hipLaunchKernelGGL((k_stag_dslash<P, PO, 8, false, false, 0, PI, FBT>), dim3(nb), dim3(256), 0, st, A);
    return;
  }
  if constexpr (NT == 16) {
    if (A.lsign) {
#define QAMD_LRG(RG)                                                                                       \
  do {                                                                                                     \
    if (red)                                                                                               \
      hipLaunchKernelGGL((k_stag_dslash<P, PO, NT, true, true, 0, PI, 0, RG>), dim3(nb), dim3(256), 0, st, A);  \
    else                                                                                                   \
      hipLaunchKernelGGL((k_stag_dslash<P, PO, NT, false, true, 0, PI, 0, RG>), dim3(nb), dim3(256), 0, st, A); \
  } while (0)
      constexpr int NG = 8 / LinkG<typename Prec<P>::S>::G;   // 3-hop groups
      switch (mall_groups<typename Prec<P>::S>(A.g)) {          // at most NG
        case 4:
          if constexpr (NG >= 4) {
            QAMD_LRG(4);
            break;
          }
          [[fallthrough]];
        case 2: QAMD_LRG(2); break;
        case 1: QAMD_LRG(1); break;
        default: QAMD_LRG(0);
      }
#undef QAMD_LRG
      return;
    }
  }
  if (red)
    hipLaunchKernelGGL((k_stag_dslash<P, PO, NT, true, false, 0, PI>), dim3(nb), dim3(256), 0, st, A);
  else
    hipLaunchKernelGGL((k_stag_dslash<P, PO, NT, false, false, 0, PI>), dim3(nb), dim3(256), 0, st, A);
}

template <int P, int PO = P, int PI = P>
static int launch_p(const qamd_links *l, Field &out, const Field &in, double a, const Field *w,
                    double b, double *red_out, hipStream_t st, const int *skip, int pset, bool finish,
                    const RedScratch *rs_fixed) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PO>::S SO;
  typedef typename Prec<PI>::S SI;
  StagArgs<P, PO, PI> A;
  memset(&A, 0, sizeof A);
  A.skip = skip;
  A.g = l->lat->g;
  A.links = (const S *)l->data[out.parity];
  A.in = (const SI *)in.data;
  A.in_norm = in.norm;
  A.out = (SO *)out.data;
  A.out_norm = out.norm;
  A.w = w ? (const SO *)w->data : nullptr;
  A.w_norm = w ? w->norm : nullptr;
  A.a = a;
  A.b = b;
  A.par = out.parity;
  A.lsign = (l->recon && (l->nterms == 16 || l->kind == 2)) ? l->lsign[out.parity] : nullptr;
  if (l->kind == 2) {
    A.links_o = (const S *)l->data[1 - out.parity];
    A.lsign_o = A.lsign ? l->lsign[1 - out.parity] : nullptr;
  }
  // Q16 links are read as raw integers: fold 1/32767 into the term scale
  for (int t = 0; t < 16; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  const bool red = red_out != nullptr || !finish;
  const int nb = (int)((A.g.vh + 255) / 256);
  if (red) {
    if (rs_fixed) {
      A.rs = *rs_fixed;
    } else {
      if (red_reserve(nb)) return -1;
      A.rs = red_set(pset);
    }
    A.red_out = red_out;
  }
  if (l->kind == 2) {
    if (l->nterms == 8) launch_t<P, PO, 8, PI, 8>(A, red, st);
    else launch_t<P, PO, 8, PI, 4>(A, red, st);
  } else if (l->nterms == 16) {
    launch_t<P, PO, 16, PI>(A, red, st);
  } else {
    launch_t<P, PO, 8, PI>(A, red, st);
  }
  if (red && finish) launch_finish<2>(A.rs, nb, FinStore2{red_out}, st, skip);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int stag_dslash_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w,
                       double b, double *red_out, hipStream_t st, const int *skip, int pset, bool finish,
                       const RedScratch *rs_fixed) {
  if (l->kind != 0 && l->kind != 2) {
    set_error("stag_dslash: links are not staggered");
    return -1;
  }
  if (pset < 0 || pset >= NSETS) {
    set_error("stag_dslash: bad partial-sum set %d", pset);
    return -1;
  }
  const bool half = (l->prec == QAMD_F16 || l->prec == QAMD_Q16);
  bool half_to_single = half && out.prec == QAMD_F32;
  // FP32 input with 16-bit links (output FP32 or in the links' format)
  const bool single_in = half && in.prec == QAMD_F32;
  if ((in.prec != l->prec && !single_in) || (out.prec != l->prec && !half_to_single) ||
      (w && w->prec != out.prec)) {
    set_error("stag_dslash: field precisions in %d / out %d do not fit links %d", in.prec, out.prec, l->prec);
    return -1;
  }
  if (out.ncomp != 3 || in.ncomp != 3 || out.nrhs != 1 || in.nrhs != 1) {
    set_error("stag_dslash: needs single-rhs colour vectors");
    return -1;
  }
  if (out.parity == in.parity) {
    set_error("stag_dslash: input and output must have opposite parity");
    return -1;
  }
  if (w && w->parity != out.parity) {
    set_error("stag_dslash: w must have the output parity");
    return -1;
  }
  switch (l->prec) {
    case QAMD_F64: return launch_p<QAMD_F64>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
    case QAMD_F32: return launch_p<QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
    case QAMD_F16:
      if (single_in && half_to_single)
        return launch_p<QAMD_F16, QAMD_F32, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish,
                                                      rs_fixed);
      if (single_in)
        return launch_p<QAMD_F16, QAMD_F16, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish,
                                                      rs_fixed);
      if (half_to_single)
        return launch_p<QAMD_F16, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
      return launch_p<QAMD_F16>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
    default:
      if (single_in && half_to_single)
        return launch_p<QAMD_Q16, QAMD_F32, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish,
                                                      rs_fixed);
      if (single_in)
        return launch_p<QAMD_Q16, QAMD_Q16, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish,
                                                      rs_fixed);
      if (half_to_single)
        return launch_p<QAMD_Q16, QAMD_F32>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
      return launch_p<QAMD_Q16>(l, out, in, a, w, b, red_out, st, skip, pset, finish, rs_fixed);
  }
}

// single-reduction CG step: t the D_oe output (opposite parity; FP32, or
// the links' 16-bit format); r, p, s, x FP32 of the output parity; links
// FP32, FP16 or Q16
template <int P, int PI>
static int cgcg_p(const qamd_links *l, const Field &t, Field &r, Field &p, Field &s, Field &x, double a,
                  CGScal *cs, int ph, RedScratch rin, RedScratch tin, RedScratch rout, hipStream_t st,
                  const int *skip) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PI>::S SI;
  StagArgs<P, QAMD_F32, PI> A;
  memset(&A, 0, sizeof A);
  A.skip = skip;
  A.g = l->lat->g;
  A.links = (const S *)l->data[r.parity];
  A.in = (const SI *)t.data;
  A.in_norm = t.norm;
  A.a = a;
  A.b = -1.0;
  A.par = r.parity;
  A.lsign = l->recon ? l->lsign[r.parity] : nullptr;
  if (l->kind == 2) {
    A.links_o = (const S *)l->data[1 - r.parity];
    A.lsign_o = A.lsign ? l->lsign[1 - r.parity] : nullptr;
  }
  for (int k = 0; k < 16; k++) A.lscale[k] = (P == QAMD_Q16) ? l->scale[k] / Q16_ONE : l->scale[k];
  A.cgr = (float2 *)r.data;
  A.cgp = (float2 *)p.data;
  A.cgs = (float2 *)s.data;
  A.cgx = (float2 *)x.data;
  const int nb = (int)((A.g.vh + 255) / 256);
  A.cg_state = cs;
  A.cg_ph = ph;
  A.cg_nb = nb;
  A.cg_rin = rin;
  A.cg_tin = tin;
  A.rs = rout;
  if (l->kind == 2) {
    if (A.lsign)
      hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 8, true, true, 2, PI, 8>), dim3(nb), dim3(256), 0, st, A);
    else
      hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 8, true, false, 2, PI, 8>), dim3(nb), dim3(256), 0, st, A);
  } else if (A.lsign) {
    constexpr int NG = 8 / LinkG<S>::G;   // 3-hop groups
    switch (mall_groups<S>(A.g)) {
      case 4:
        if constexpr (NG >= 4) {
          hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 16, true, true, 2, PI, 0, 4>), dim3(nb), dim3(256), 0, st, A);
          break;
        }
        [[fallthrough]];
      case 2:
        hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 16, true, true, 2, PI, 0, 2>), dim3(nb), dim3(256), 0, st, A);
        break;
      case 1:
        hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 16, true, true, 2, PI, 0, 1>), dim3(nb), dim3(256), 0, st, A);
        break;
      default: hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 16, true, true, 2, PI>), dim3(nb), dim3(256), 0, st, A);
    }
  } else {
    hipLaunchKernelGGL((k_stag_dslash<P, QAMD_F32, 16, true, false, 2, PI>), dim3(nb), dim3(256), 0, st, A);
  }
  QAMD_HIP(hipGetLastError());
  return 0;
}

int stag_cgcg_step_launch(const qamd_links *l, const Field &t, Field &r, Field &p, Field &s, Field &x, double a,
                          CGScal *cs, int ph, RedScratch rin, RedScratch tin, RedScratch rout, hipStream_t st,
                          const int *skip) {
  if (!((l->kind == 0 && l->nterms == 16) || (l->kind == 2 && l->nterms == 8)) || l->prec == QAMD_F64 ||
      (t.prec != QAMD_F32 && t.prec != l->prec) || t.parity == r.parity ||
      r.prec != QAMD_F32 || p.prec != QAMD_F32 || s.prec != QAMD_F32 || x.prec != QAMD_F32 ||
      p.parity != r.parity || s.parity != r.parity || x.parity != r.parity || t.nrhs != 1 || r.nrhs != 1 ||
      (ph != 0 && ph != 1)) {
    set_error("stag_cgcg_step: needs sloppy asqtad links (16 terms or forward-only fat + long) and single-rhs "
              "FP32 t (opposite parity) and r, p, s, x");
    return -1;
  }
  switch (l->prec) {
    case QAMD_F32: return cgcg_p<QAMD_F32, QAMD_F32>(l, t, r, p, s, x, a, cs, ph, rin, tin, rout, st, skip);
    case QAMD_F16:
      if (t.prec == QAMD_F16) return cgcg_p<QAMD_F16, QAMD_F16>(l, t, r, p, s, x, a, cs, ph, rin, tin, rout, st, skip);
      return cgcg_p<QAMD_F16, QAMD_F32>(l, t, r, p, s, x, a, cs, ph, rin, tin, rout, st, skip);
    default:
      if (t.prec == QAMD_Q16) return cgcg_p<QAMD_Q16, QAMD_Q16>(l, t, r, p, s, x, a, cs, ph, rin, tin, rout, st, skip);
      return cgcg_p<QAMD_Q16, QAMD_F32>(l, t, r, p, s, x, a, cs, ph, rin, tin, rout, st, skip);
  }
}

// ------------------------------------------------- no-allocation variant
// Partial sums and totals of qamd_dslash_noalloc live in this fixed array
// (one copy per GPU, held until the process ends): row i of the partial
// sums at [i * NOALLOC_BLOCKS + block], the two totals at 2 * NOALLOC_BLOCKS.
// Every launch writes each partial and total it later reads, so nothing has
// to start at zero.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
constexpr int NOALLOC_BLOCKS = 65536;   // 256-site blocks: up to 2^24 sites per parity
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS + 8];
static double *g_scratch_dev[64];       // per device ordinal

static int noalloc_scratch(double **ptr) {
  int dev = 0;
  QAMD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("dslash_noalloc: device ordinal %d out of range", dev);
    return -1;
  }
  if (!g_scratch_dev[dev]) {
    void *q = nullptr;
    QAMD_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = (double *)q;
  }
  *ptr = g_scratch_dev[dev];
  return 0;
}

}  // namespace qamd

using namespace qamd;

extern "C" {

int qamd_dslash_noalloc(const qamd_links *l, qamd_field *out, const qamd_field *in, double a,
                        const qamd_field *w, double b, double *norm2_out) {
  if (!l || !out || !in) {
    set_error("dslash_noalloc: NULL argument");
    return -1;
  }
  if (!initialized()) {
    set_error("dslash_noalloc: call qamd_init first (this entry point allocates nothing)");
    return -1;
  }
  if ((l->kind != 0 && l->kind != 2) || in->f.nrhs != 1 || out->f.nrhs != 1) {
    set_error("dslash_noalloc: needs staggered links and single-rhs fields");
    return -1;
  }
  const int64_t nb = (l->lat->g.vh + 255) / 256;
  if (nb > NOALLOC_BLOCKS) {
    set_error("dslash_noalloc: %lld sites per parity exceed the scratch limit of %lld", (long long)l->lat->g.vh,
              (long long)NOALLOC_BLOCKS * 256);
    return -1;
  }
  double *s = nullptr;
  if (norm2_out && noalloc_scratch(&s)) return -1;
  const RedScratch rs{s, NOALLOC_BLOCKS};
  double *tot = s ? s + 2 * NOALLOC_BLOCKS : nullptr;
  if (stag_dslash_launch(l, out->f, in->f, a, w ? &w->f : nullptr, b, tot, stream(), nullptr, 0, true,
                         norm2_out ? &rs : nullptr))
    return -1;
  if (norm2_out) {
    double t[2];
    QAMD_HIP(hipMemcpyAsync(t, tot, sizeof t, hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    *norm2_out = t[0];
  }
  return 0;
}

int qamd_dslash(const qamd_links *l, qamd_field *out, const qamd_field *in, double a,
                const qamd_field *w, double b, double *norm2_out) {
  if (!l || !out || !in) {
    set_error("dslash: NULL argument");
    return -1;
  }
  double *slot = nullptr;
  if (norm2_out) {
    slot = scalar_slot();
    if (!slot) return -1;
  }
  if (l->kind == 0 && in->f.nrhs > 1) {   // block of rhs: |out|^2 summed over the block
    if (stag_block_launch(l, out->f, in->f, a, w ? &w->f : nullptr, b, false, RedScratch{}, stream())) return -1;
    if (norm2_out) {
      double t[16];
      if (qamd_field_norm2(out, t)) return -1;
      *norm2_out = 0;
      for (int r = 0; r < out->f.nrhs; r++) *norm2_out += t[r];
    }
    return 0;
  }
  int rc = (l->kind == 1)
               ? wilson_dslash_launch(l, out->f, in->f, a, w ? &w->f : nullptr, b, slot, stream(), nullptr)
               : stag_dslash_launch(l, out->f, in->f, a, w ? &w->f : nullptr, b, slot, stream(), nullptr);
  if (rc) return -1;
  if (norm2_out) {
    double t[2];
    QAMD_HIP(hipMemcpyAsync(t, slot, sizeof t, hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    *norm2_out = t[0];
  }
  return 0;
}

int qamd_stag_normal(const qamd_links *l, qamd_field *Ap, const qamd_field *p, qamd_field *tmp,
                     double mass, double *pAp) {
  if (p->f.nrhs > 1) {   // a block: pAp summed over the rhs
    if (stag_block_launch(l, tmp->f, p->f, 0.0, nullptr, 1.0, false, RedScratch{}, stream())) return -1;
    if (stag_block_launch(l, Ap->f, tmp->f, mass * mass, &p->f, -1.0, false, RedScratch{}, stream())) return -1;
    if (pAp) {
      double d[16];
      if (qamd_field_redot(p, Ap, d)) return -1;
      *pAp = 0;
      for (int r = 0; r < p->f.nrhs; r++) *pAp += d[r];
    }
    return 0;
  }
  double *slot = scalar_slot();
  if (!slot) return -1;
  // tmp = D_oe p  (|tmp|^2 -> slot[0])
  if (stag_dslash_launch(l, tmp->f, p->f, 0.0, nullptr, 1.0, slot, stream(), nullptr)) return -1;
  // Ap = m^2 p - D_eo tmp  (|p|^2 -> slot[3])
  if (stag_dslash_launch(l, Ap->f, tmp->f, mass * mass, &p->f, -1.0, slot + 2, stream(), nullptr)) return -1;
  if (pAp) {
    double t[4];
    QAMD_HIP(hipMemcpyAsync(t, slot, sizeof t, hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    *pAp = mass * mass * t[3] + t[0];
  }
  return 0;
}

int qamd_set_mall_share(int mb64, int mb) {
  g_mall_mb[0] = mb64 < 0 ? -1 : mb64;
  g_mall_mb[1] = mb < 0 ? -1 : mb;
  return 0;
}

int qamd_time_dslash(const qamd_links *l, qamd_field *out, const qamd_field *in, int nrep,
                     double *ms_per_call) {
  hipEvent_t e0, e1;
  QAMD_HIP(hipEventCreate(&e0));
  QAMD_HIP(hipEventCreate(&e1));
  QAMD_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < nrep; i++) {
    int rc = (l->kind == 1)   ? wilson_dslash_launch(l, out->f, in->f, 0.0, nullptr, 1.0, nullptr, stream())
             : in->f.nrhs > 1 ? stag_block_launch(l, out->f, in->f, 0.0, nullptr, 1.0, false, RedScratch{}, stream())
                              : stag_dslash_launch(l, out->f, in->f, 0.0, nullptr, 1.0, nullptr, stream(), nullptr);
    if (rc) return -1;
  }
  QAMD_HIP(hipEventRecord(e1, stream()));
  QAMD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_call = ms / nrep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
