// Block-of-rhs staggered Dslash, shared by the single-problem kernel
// (stag_block.hip) and the fused multi-problem batch kernels (batch.hip).
#pragma once
#include "qamd_internal.h"
#include "qamd_vec3.h"

namespace qamd {

// links in precision P, input in PI (P, or FP32 with 16-bit links),
// output and w in PO
template <int P, int PO, int PI = P> struct MArgs {
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
  double ar[16], b;        // slot q: out = ar[q]*w + b * D in, for rhs rmap[q]
  unsigned char rmap[16];  // rhs index of slot q < nr (the rhs a launch computes)
  // fused block-CG step (CGX, FP32 vectors): with w = p, Ap = ar p + b D in
  // is not stored; x_r += alpha[r] p_r, r_r -= alpha[r] Ap_r, |r_r|^2 -> red[2q]
  float2 *cgx, *cgr;
  const double *alpha;
  int par, nr;
  float lscale[16];
  RedScratch rs;   // partial sums of rhs r in slots 2r, 2r+1 (offset by the caller)
  int xcd;         // block order: 0 plain, 1 XCD ranges (xcd_block), 2 slab sweep (slab_block)
  SlabOrder slab;
  const uint8_t *lsign;   // set: long terms' third rows rebuilt (determinant sign bits per site)
};

// Slots per step of the term loop: each step loads the links of SPS
// consecutive slots (in group_at order), then runs the rhs loop over them, so
// the gathers of one rhs from those slots' neighbours -- the 1- and 3-hop
// neighbours along one direction, which share cache lines -- are issued back
// to back instead of one slot (all rhs) apart.  With eight rhs a wavefront's
// slots are several microseconds apart and the shared lines are gone from L2
// by the next slot (32^4, 8 rhs, one MI355X: FP64 0.58 -> 0.43 ms with 2
// slots per step, FP32 0.27 -> 0.20 ms with 4, Q16 0.197 -> 0.180 ms with one
// 4-slot group).  Every rhs still adds its terms in slot order, so the sums
// are the same bits for any step size.
template <int P, int NR> struct BlockStep {
  static constexpr int G = LinkG<typename Prec<P>::S>::G;
  static constexpr int SPS = P == QAMD_F64 ? (NR >= 2 ? 2 : 1) : P == QAMD_F32 ? (NR >= 2 && NR <= 8 ? 4 : 2) : G;
};

// one output site cb: for the slots q < A.nr, rhs r = A.rmap[q]:
// out_r = ar[q]*w_r + b * sum_t L_t in_r(x+d_t) (fields rhs-outer); with
// RED, red[2q] += |out_r|^2, red[2q+1] += |w_r|^2
template <int P, int PO, int NT, int NR, bool RED, bool CGX = false, int PI = P>
__device__ __forceinline__ void block_site(const MArgs<P, PO, PI> &A, int cb, double (&red)[2 * NR]) {
  static_assert(!CGX || (PO == QAMD_F32 && RED), "fused block-CG step: FP32 vectors, with the |r|^2 sums");
  static_assert(PI == P || (PI == QAMD_F32 && (P == QAMD_F16 || P == QAMD_Q16)),
                "input precision: the links' or FP32 with 16-bit links");
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::S S;
  const int64_t vp = A.g.vh_pad;
  const int nr = A.nr;
  Site s = site_of(A.g, cb, A.par);
  constexpr int G = LinkG<S>::G;
  constexpr int SPS = BlockStep<P, NR>::SPS;
  constexpr int GPS = SPS / G;   // slot groups per step
  static_assert(SPS % G == 0 && (NT / G) % GPS == 0, "a step is whole slot groups");
  const S *lp = link_base<S, NT>(A.links, cb);
  const S *lp3 = link_base3<S, NT>(lp, vp);   // the 3-hop terms' region
  C acc[NR][3];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int j = 0; j < 3; j++) acc[r][j].x = acc[r][j].y = (R)0;
  // the long terms' third rows are rebuilt when the link set allows it (a
  // per-problem, hence workgroup-uniform, choice)
  const bool rec = NT == 16 && A.lsign != nullptr;
  const unsigned lsb = rec ? A.lsign[cb] : 0u;
#pragma unroll
  for (int st = 0; st < NT / G / GPS; st++) {
    C m[SPS][9];
    int nb[SPS];
    R sc[SPS];
#pragma unroll
    for (int gi = 0; gi < GPS; gi++) {
      // one 16-byte group load per element serves the G slots of group sg
      const int sg = group_at<G, NT>(st * GPS + gi);
      const bool lng = NT == 16 && sg * G >= 8;
      S q[9][G];
#pragma unroll
      for (int e = 0; e < 9; e++)
        if (!(lng && rec && e % 3 == 2)) ntload_group((lng ? lp3 : lp) + link_off<S, NT>(sg * G, e), q[e]);
#pragma unroll
      for (int gg = 0; gg < G; gg++) {
        const int j = gi * G + gg;
        const int sl = sg * G + gg;
        const int t = slot_term(sl, NT);
        const int k = (NT == 16) ? t / 2 : t;
        const int hop = (NT == 16 && (t & 1)) ? 3 : 1;
        const int mu = k / 2;
        const int h = (k & 1) ? -hop : hop;
        nb[j] = nbr_cb(A.g, s, mu, h);
        sc[j] = (P == QAMD_F16 || P == QAMD_Q16) ? A.lscale[t] : (R)1;
#pragma unroll
        for (int e = 0; e < 9; e++)
          if (!(lng && rec && e % 3 == 2)) m[j][e] = to_c(q[e][gg], (R)0);
        if (lng && rec) long_row3<C, R>(m[j], (lsb >> (sl - 8)) & 1);
        if constexpr (PI != P) {   // FP32 input: the term's link scale applied to the link, once for all rhs
#pragma unroll
          for (int e = 0; e < 9; e++) {
            m[j][e].x *= sc[j];
            m[j][e].y *= sc[j];
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (r < nr) {
        const int64_t ri = A.rmap[r];
#pragma unroll
        for (int j = 0; j < SPS; j++) {
          C v[3];
          if constexpr (PI == P)
            VecLd<P>::ld(A.in + ri * 3 * vp, A.in_norm ? A.in_norm + ri * vp : nullptr, vp, nb[j], v, sc[j]);
          else
            VecLd<PI>::ld(A.in + ri * 3 * vp, nullptr, vp, nb[j], v, (R)1);
#pragma unroll
          for (int row = 0; row < 3; row++)
#pragma unroll
            for (int c = 0; c < 3; c++) cmac<C, R>(acc[r][row], m[j][3 * c + row], v[c]);
        }
      }
    }
  }
  const R b = (R)A.b;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    if (r < nr) {
      const int64_t rr = A.rmap[r];
      const R a = (R)A.ar[r];
      C o[3];
      if constexpr (CGX) {
        C w[3];   // p_rr
        VecLd<PO>::ld(A.w + rr * 3 * vp, nullptr, vp, cb, w, (R)1);
        const R al = (R)A.alpha[rr];
        float2 xv[3], rv[3];
        VecLd<QAMD_F32>::ld(A.cgx + rr * 3 * vp, nullptr, vp, cb, xv, 1.0f);
        VecLd<QAMD_F32>::ld(A.cgr + rr * 3 * vp, nullptr, vp, cb, rv, 1.0f);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          o[j].x = fma(a, w[j].x, b * acc[r][j].x);   // Ap
          o[j].y = fma(a, w[j].y, b * acc[r][j].y);
          xv[j].x = fma(al, w[j].x, xv[j].x);
          xv[j].y = fma(al, w[j].y, xv[j].y);
          rv[j].x = fma(-al, o[j].x, rv[j].x);
          rv[j].y = fma(-al, o[j].y, rv[j].y);
          red[2 * r] += (double)rv[j].x * rv[j].x + (double)rv[j].y * rv[j].y;
        }
        VecSt<QAMD_F32>::st(A.cgx + rr * 3 * vp, nullptr, vp, cb, xv);
        VecSt<QAMD_F32>::st(A.cgr + rr * 3 * vp, nullptr, vp, cb, rv);
        continue;
      }
      if (A.w) {
        C w[3];
        VecLd<PO>::ld(A.w + rr * 3 * vp, A.w_norm ? A.w_norm + rr * vp : nullptr, vp, cb, w, (R)1);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          o[j].x = fma(a, w[j].x, b * acc[r][j].x);
          o[j].y = fma(a, w[j].y, b * acc[r][j].y);
          if (RED) red[2 * r + 1] += (double)w[j].x * w[j].x + (double)w[j].y * w[j].y;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          o[j].x = b * acc[r][j].x;
          o[j].y = b * acc[r][j].y;
        }
      }
      if (RED)
#pragma unroll
        for (int j = 0; j < 3; j++) red[2 * r] += (double)o[j].x * o[j].x + (double)o[j].y * o[j].y;
      VecSt<PO>::st(A.out + rr * 3 * vp, A.out_norm ? A.out_norm + rr * vp : nullptr, vp, cb, o);
    }
  }
}

}  // namespace qamd
