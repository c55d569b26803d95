// Blocks of right-hand sides for the staggered operator: a Dslash that loads
// every link once for up to 16 right-hand sides (thread per site, the rhs
// loop innermost, so the link stream is divided by the block size), and a
// mixed-precision block CG with per-rhs scalars and reliable updates.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"
#include "qamd_vec3.h"
#include "qamd_block.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <chrono>
#include <stdio.h>

namespace qamd {

constexpr int MAXB = 16;

// out_r = a*w_r + b * sum_t L_t in_r(x+d_t) for the rhs r = rmap[q], q < nr;
// fields rhs-outer
template <int P, int PO, int NT, int NR, bool RED, bool CGX = false, int PI = P>
__global__ void __launch_bounds__(256) k_stag_block(MArgs<P, PO, PI> A) {
  const int blk = A.xcd == 2 ? slab_block(blockIdx.x, A.slab) : A.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  int cb = blk * 256 + threadIdx.x;
  double red[2 * NR];
#pragma unroll
  for (int i = 0; i < 2 * NR; i++) red[i] = 0.0;
  if (cb < A.g.vh) block_site<P, PO, NT, NR, RED, CGX, PI>(A, cb, red);
  // slots of rhs r at 2r, 2r+1 (fused CG step: |r_r|^2 at slot r), indexed by
  // site block, so the sums do not depend on the block order or on which
  // launch computed the rhs
  if (RED) {
    block_sum(red);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < NR; q++)
        if (q < A.nr) {
          if (CGX) {
            A.rs.part[A.rmap[q] * A.rs.max_blocks + blk] = red[2 * q];
          } else {
            A.rs.part[(2 * A.rmap[q]) * A.rs.max_blocks + blk] = red[2 * q];
            A.rs.part[(2 * A.rmap[q] + 1) * A.rs.max_blocks + blk] = red[2 * q + 1];
          }
        }
  }
}

template <int P, int PO, int NT, int NR, int PI>
static void blaunch_nr(const MArgs<P, PO, PI> &A, bool red, hipStream_t st) {
  int nb = (int)((A.g.vh + 255) / 256);
  if (red) hipLaunchKernelGGL((k_stag_block<P, PO, NT, NR, true, false, PI>), dim3(nb), dim3(256), 0, st, A);
  else hipLaunchKernelGGL((k_stag_block<P, PO, NT, NR, false, false, PI>), dim3(nb), dim3(256), 0, st, A);
}

template <int P, int PO, int NT, int PI>
static int blaunch_nt(const MArgs<P, PO, PI> &A, bool red, hipStream_t st) {
  if (A.nr <= 1) blaunch_nr<P, PO, NT, 1>(A, red, st);
  else if (A.nr <= 2) blaunch_nr<P, PO, NT, 2>(A, red, st);
  else if (A.nr <= 4) blaunch_nr<P, PO, NT, 4>(A, red, st);
  else if (A.nr <= 8) blaunch_nr<P, PO, NT, 8>(A, red, st);
  else if constexpr (P != QAMD_F64) blaunch_nr<P, PO, NT, 16>(A, red, st);
  else {
    set_error("block Dslash: FP64 blocks hold at most 8 rhs");
    return -1;
  }
  return 0;
}

// One block-Dslash application, in launches of k rhs with or without the
// XCD block order; (k, xcd) is chosen by the autotuner per lattice,
// precision and block size.
template <int P, int PO, int PI>
struct BlockRun {
  MArgs<P, PO, PI> A;
  int n, nt, red, ncand, k[10], xcd[10];
  hipStream_t st;
  static int run(int c, void *ctx) {
    BlockRun &R = *(BlockRun *)ctx;
    for (int r0 = 0; r0 < R.n; r0 += R.k[c]) {
      MArgs<P, PO, PI> B = R.A;   // slots r0 .. r0+nr-1 of the block's slots
      B.nr = R.n - r0 < R.k[c] ? R.n - r0 : R.k[c];
      for (int q = 0; q < B.nr; q++) {
        B.ar[q] = R.A.ar[r0 + q];
        B.rmap[q] = R.A.rmap[r0 + q];
      }
      B.xcd = R.xcd[c];
      int rc = (R.nt == 16) ? blaunch_nt<P, PO, 16, PI>(B, R.red, R.st) : blaunch_nt<P, PO, 8, PI>(B, R.red, R.st);
      if (rc) return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
};

template <int P, int PO = P, int PI = P>
static int blaunch(const qamd_links *l, Field &out, const Field &in, const double *ar, const Field *w, double b,
                   bool red, RedScratch rs, hipStream_t st, unsigned act) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PO>::S SO;
  typedef typename Prec<PI>::S SI;
  BlockRun<P, PO, PI> R;
  memset(&R, 0, sizeof R);
  MArgs<P, PO, PI> &A = R.A;
  A.g = l->lat->g;
  A.links = (const S *)l->data[out.parity];
  A.in = (const SI *)in.data;
  A.in_norm = in.norm;
  A.out = (SO *)out.data;
  A.out_norm = out.norm;
  A.w = w ? (const SO *)w->data : nullptr;
  A.w_norm = w ? w->norm : nullptr;
  // the rhs to compute, in order, as the launch's slots
  int ns = 0;
  for (int r = 0; r < in.nrhs; r++)
    if ((act >> r) & 1u) {
      A.ar[ns] = ar[r];
      A.rmap[ns++] = (unsigned char)r;
    }
  A.b = b;
  A.par = out.parity;
  A.nr = ns;
  A.rs = rs;
  A.xcd = 1;
  A.lsign = (l->recon && l->nterms == 16) ? l->lsign[out.parity] : nullptr;
  for (int t = 0; t < 16; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  if (ns == 0) return 0;   // no rhs to compute
  R.n = ns;
  R.nt = l->nterms;
  R.red = red;
  R.st = st;
  // candidates: all rhs in one launch (default), then halves, quarters, ...,
  // each with the XCD block order and without; all rhs in one launch in the
  // slab-sweep order when the lattice allows it
  const bool slab = slab_order(A.g, A.slab);
  for (int kk = R.n; R.ncand < 8; kk = (kk + 1) / 2) {
    R.k[R.ncand] = kk;
    R.xcd[R.ncand++] = 1;
    R.k[R.ncand] = kk;
    R.xcd[R.ncand++] = 0;
    if (kk == 1) break;
  }
  if (slab) {
    R.k[R.ncand] = R.n;
    R.xcd[R.ncand++] = 2;
  }
  if (const char *o = getenv("QAMD_BLOCK_ORDER")) {   // force an order (measurements)
    R.ncand = 1;
    R.k[0] = R.n;
    R.xcd[0] = (atoi(o) == 2 && !slab) ? 1 : atoi(o);
  }
  // out = a*out + ... would compound over the timed repeats: no tuning then;
  // nor for launches over a subset of the rhs (solvers whose converged rhs
  // drop out): they take the default shape rather than tune once per count
  const bool inplace = (w && w->data == out.data) || ns != in.nrhs;
  char key[128];
  snprintf(key, sizeof key, "stag_block p%d/%d/%d nt%d n%d red%d %dx%dx%dx%d", P, PO, PI, R.nt, R.n, (int)red,
           A.g.L[0], A.g.L[1], A.g.L[2], A.g.L[3]);
  int c = inplace ? 0 : tune_pick(key, R.ncand, 0, &BlockRun<P, PO, PI>::run, &R);
  if (BlockRun<P, PO, PI>::run(c, &R)) {
    set_error("block Dslash: launch failed");
    return -1;
  }
  return 0;
}

// block Dslash out_r = ar[r] w_r + b D in_r for the rhs r whose bit is set
// in act; with red, 2*nrhs partial sums {|out_r|^2, |w_r|^2} go to rs
int stag_block_launch_m(const qamd_links *l, Field &out, const Field &in, const double *ar, const Field *w,
                        double b, bool red, RedScratch rs, hipStream_t st, unsigned act) {
  const bool half = (l->prec == QAMD_F16 || l->prec == QAMD_Q16);
  const bool h2s = half && out.prec == QAMD_F32;
  // FP32 input with 16-bit links (output FP32 or in the links' format)
  const bool single_in = half && in.prec == QAMD_F32;
  if (l->kind != 0 || (in.prec != l->prec && !single_in) || (out.prec != l->prec && !h2s) ||
      (w && w->prec != out.prec)) {
    set_error("block Dslash: precision/kind mismatch");
    return -1;
  }
  if (in.nrhs != out.nrhs || (w && w->nrhs != in.nrhs) || in.nrhs > MAXB || in.ncomp != 3 || out.ncomp != 3 ||
      in.rhs_inner || out.rhs_inner || out.parity == in.parity || (w && w->parity != out.parity)) {
    set_error("block Dslash: field shapes do not match");
    return -1;
  }
  if (red && red_reserve((int)((l->lat->g.vh + 255) / 256))) return -1;
  switch (l->prec) {
    case QAMD_F64: return blaunch<QAMD_F64>(l, out, in, ar, w, b, red, rs, st, act);
    case QAMD_F32: return blaunch<QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act);
    case QAMD_F16:
      if (single_in)
        return h2s ? blaunch<QAMD_F16, QAMD_F32, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act)
                   : blaunch<QAMD_F16, QAMD_F16, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act);
      return h2s ? blaunch<QAMD_F16, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act)
                 : blaunch<QAMD_F16>(l, out, in, ar, w, b, red, rs, st, act);
    default:
      if (single_in)
        return h2s ? blaunch<QAMD_Q16, QAMD_F32, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act)
                   : blaunch<QAMD_Q16, QAMD_Q16, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act);
      return h2s ? blaunch<QAMD_Q16, QAMD_F32>(l, out, in, ar, w, b, red, rs, st, act)
                 : blaunch<QAMD_Q16>(l, out, in, ar, w, b, red, rs, st, act);
  }
}

// Fused step of the block CG: for the rhs set in act, Ap_r = m2[r] p_r -
// D in_r (FP32, not stored), x_r += alpha[r] p_r, r_r -= alpha[r] Ap_r, and
// |r_r|^2 partial sums at slot r of rs.  All rhs in one launch in the XCD
// block order (no tuning: trial launches would repeat the in-place update).
template <int P, int PI = P>
static int bcgx(const qamd_links *l, const Field &in, const Field &p, Field &x, Field &r, const double *m2,
                const double *alpha, RedScratch rs, hipStream_t st, unsigned act) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PI>::S SI;
  MArgs<P, QAMD_F32, PI> A;
  memset(&A, 0, sizeof A);
  A.g = l->lat->g;
  A.links = (const S *)l->data[r.parity];
  A.in = (const SI *)in.data;
  A.in_norm = in.norm;
  A.w = (const float2 *)p.data;
  A.b = -1.0;
  A.par = r.parity;
  A.rs = rs;
  A.xcd = 1;
  A.cgx = (float2 *)x.data;
  A.cgr = (float2 *)r.data;
  A.alpha = alpha;
  A.lsign = (l->recon && l->nterms == 16) ? l->lsign[r.parity] : nullptr;
  for (int t = 0; t < 16; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  int ns = 0;
  for (int q = 0; q < p.nrhs; q++)
    if ((act >> q) & 1u) {
      A.ar[ns] = m2[q];
      A.rmap[ns++] = (unsigned char)q;
    }
  A.nr = ns;
  if (ns == 0) return 0;
  const int nb = (int)((A.g.vh + 255) / 256);
#define CGXL(NR) hipLaunchKernelGGL((k_stag_block<P, QAMD_F32, 16, NR, true, true, PI>), dim3(nb), dim3(256), 0, st, A)
  if (ns <= 1) CGXL(1);
  else if (ns <= 2) CGXL(2);
  else if (ns <= 4) CGXL(4);
  else if (ns <= 8) CGXL(8);
  else CGXL(16);
#undef CGXL
  QAMD_HIP(hipGetLastError());
  return 0;
}

static int stag_block_cg_step(const qamd_links *l, const Field &in, const Field &p, Field &x, Field &r,
                              const double *m2, const double *alpha, RedScratch rs, hipStream_t st, unsigned act) {
  const bool single_in = l->prec != QAMD_F32 && in.prec == QAMD_F32;   // FP32 D_oe p with 16-bit links
  if (l->kind != 0 || l->nterms != 16 || l->prec == QAMD_F64 || (in.prec != l->prec && !single_in) ||
      p.prec != QAMD_F32 ||
      x.prec != QAMD_F32 || r.prec != QAMD_F32 || p.nrhs > MAXB || in.nrhs != p.nrhs || x.nrhs != p.nrhs ||
      r.nrhs != p.nrhs) {
    set_error("block CG step: needs 16-term sloppy links and FP32 p, x, r blocks");
    return -1;
  }
  if (red_reserve((int)((l->lat->g.vh + 255) / 256))) return -1;
  switch (l->prec) {
    case QAMD_F32: return bcgx<QAMD_F32>(l, in, p, x, r, m2, alpha, rs, st, act);
    case QAMD_F16:
      return single_in ? bcgx<QAMD_F16, QAMD_F32>(l, in, p, x, r, m2, alpha, rs, st, act)
                       : bcgx<QAMD_F16>(l, in, p, x, r, m2, alpha, rs, st, act);
    default:
      return single_in ? bcgx<QAMD_Q16, QAMD_F32>(l, in, p, x, r, m2, alpha, rs, st, act)
                       : bcgx<QAMD_Q16>(l, in, p, x, r, m2, alpha, rs, st, act);
  }
}

// every rhs, one coefficient a
int stag_block_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                      bool red, RedScratch rs, hipStream_t st) {
  double ar[MAXB];
  for (int r = 0; r < MAXB; r++) ar[r] = a;
  return stag_block_launch_m(l, out, in, ar, w, b, red, rs, st, ~0u);
}

// ------------------------------------------------------------ block CG
struct BatchDev {
  double rsq[MAXB], rsq_old[MAXB], t2[MAXB], p2[MAXB], beta[MAXB], maxrr[MAXB];
  int active[MAXB], reset[MAXB];
  double m2[MAXB];   // mass^2 (plus shift) of rhs r's operator
  double alpha[MAXB];
  int nrhs;
};

template <int N, class Fin>
__global__ void __launch_bounds__(FT) k_finish_b2(RedScratch a, RedScratch b, int nb, Fin fin) {
  double w0[N], w1[N], t[2 * N];
  finish_sum(a, nb, w0);
  finish_sum(b, nb, w1);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      t[i] = w0[i];
      t[N + i] = w1[i];
    }
    fin(t);
  }
}

// The epilogues loop over a compile-time block size NR (slots of rhs >= nrhs
// are skipped), so the totals stay in registers.
template <int NR> struct FinBNormal {   // set a: {|t_r|^2, 0}, set b: {|Ap_r|^2, |p_r|^2}
  BatchDev *B;
  __device__ void operator()(const double *t) const {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < B->nrhs) {
        B->t2[r] = t[2 * r];
        B->p2[r] = t[2 * NR + 2 * r + 1];
      }
  }
};

template <int NR> struct FinBCG {
  BatchDev *B;
  __device__ void operator()(const double *t) const {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < B->nrhs && B->active[r]) {
        B->rsq_old[r] = B->rsq[r];
        B->rsq[r] = t[r];
        B->beta[r] = B->rsq[r] / B->rsq_old[r];
        B->maxrr[r] = fmax(B->maxrr[r], sqrt(B->rsq[r]));
      }
  }
};

template <int NR> struct FinBTrue {   // active rhs only: the others keep their last true |r|^2
  BatchDev *B;
  __device__ void operator()(const double *t) const {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < B->nrhs && B->active[r]) {
        B->rsq[r] = t[r];
        B->maxrr[r] = sqrt(t[r]);
      }
  }
};

// alpha_r = rsq_r / (m2_r |p_r|^2 + |t_r|^2) from |t_r|^2 (slot 2r of the
// Dslash sums) and |p_r|^2 (slot r of the p-update sums), active rhs
template <int NR> struct FinBAlpha {
  BatchDev *B;
  __device__ void operator()(const double *t) const {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < B->nrhs && B->active[r]) {
        B->t2[r] = t[2 * r];
        B->p2[r] = t[2 * NR + r];
        B->alpha[r] = B->rsq[r] / (B->m2[r] * B->p2[r] + B->t2[r]);
      }
  }
};

// finishers sized to the block: NR = block size rounded up to 1, 2, 4, 8, 16
template <int NR> static void fin_normal(int nb, BatchDev *B, hipStream_t st) {
  hipLaunchKernelGGL((k_finish_b2<2 * NR, FinBNormal<NR>>), dim3(1), dim3(FT), 0, st, red_set(1), red_set(2), nb,
                     FinBNormal<NR>{B});
}
template <int NR> static void fin_cg(int nb, BatchDev *B, hipStream_t st) {
  launch_finish<NR>(red_set(0), nb, FinBCG<NR>{B}, st);
}
template <int NR> static void fin_true(int nb, BatchDev *B, hipStream_t st) {
  launch_finish<NR>(red_set(0), nb, FinBTrue<NR>{B}, st);
}
template <int NR> static void fin_alpha(int nb, BatchDev *B, hipStream_t st) {
  hipLaunchKernelGGL((k_finish2<2 * NR, NR, FinBAlpha<NR>>), dim3(1), dim3(FT), 0, st, red_set(1), red_set(3), nb,
                     FinBAlpha<NR>{B}, nullptr);
}
struct FinOps {
  void (*normal)(int, BatchDev *, hipStream_t);
  void (*cg)(int, BatchDev *, hipStream_t);
  void (*truer)(int, BatchDev *, hipStream_t);
  void (*alpha)(int, BatchDev *, hipStream_t);
};
template <int NR> static FinOps fin_ops_nr() {
  return FinOps{fin_normal<NR>, fin_cg<NR>, fin_true<NR>, fin_alpha<NR>};
}
static FinOps fin_ops(int n) {
  if (n <= 1) return fin_ops_nr<1>();
  if (n <= 2) return fin_ops_nr<2>();
  if (n <= 4) return fin_ops_nr<4>();
  if (n <= 8) return fin_ops_nr<8>();
  return fin_ops_nr<16>();
}

// per-rhs slot of the block sums: grid.y = rhs
__device__ __forceinline__ void rhs_partial(double v, RedScratch rs) {
  double a[1] = {v};
  block_sum(a);
  if (threadIdx.x == 0) rs.part[blockIdx.y * rs.max_blocks + blockIdx.x] = a[0];
}

template <int P> struct VB {
  typedef typename Prec<P>::S S;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  __device__ static void ld(const Field &f, int r, int i, C (&v)[3]) {
    VecLd<P>::ld((const S *)f.data + (int64_t)r * 3 * f.vp, f.norm ? f.norm + (int64_t)r * f.vp : nullptr, f.vp, i,
                 v, (R)1);
  }
  __device__ static void st(Field &f, int r, int i, const C (&v)[3]) {
    VecSt<P>::st((S *)f.data + (int64_t)r * 3 * f.vp, f.norm ? f.norm + (int64_t)r * f.vp : nullptr, f.vp, i, v);
  }
};

// p_r = r_r + beta_r p_r (mode 0, active rhs), p_r = r_r (mode 1: rhs
// flagged in reset; mode 2: every active rhs), and |p_r|^2 partial sums
// (slot r), which the next alpha_r needs before the fused step
__global__ void __launch_bounds__(256) k_bcg_pn(Field p, Field r, BatchDev *B, RedScratch rs, int mode) {
  const int rr = blockIdx.y;
  const bool on = mode == 1 ? B->reset[rr] != 0 : B->active[rr] != 0;
  if (!on) return;   // uniform over the block: its partial-sum slot is not read
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < p.g->vh) {
    float2 pv[3], rv[3];
    VB<QAMD_F32>::ld(r, rr, i, rv);
    if (mode == 0) {
      const float be = (float)B->beta[rr];
      VB<QAMD_F32>::ld(p, rr, i, pv);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        pv[j].x = fmaf(be, pv[j].x, rv[j].x);
        pv[j].y = fmaf(be, pv[j].y, rv[j].y);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 3; j++) pv[j] = rv[j];
    }
#pragma unroll
    for (int j = 0; j < 3; j++) red += (double)pv[j].x * pv[j].x + (double)pv[j].y * pv[j].y;
    VB<QAMD_F32>::st(p, rr, i, pv);
  }
  rhs_partial(red, rs);
}

// reliable update: x64 += xs, xs = 0 (active rhs)
__global__ void k_bru_accum(Field x, Field xs, BatchDev *B) {
  const int rr = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!B->active[rr] || i >= x.g->vh) return;
  double2 h[3];
  float2 s[3];
  VB<QAMD_F64>::ld(x, rr, i, h);
  VB<QAMD_F32>::ld(xs, rr, i, s);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    h[j].x += s[j].x;
    h[j].y += s[j].y;
    s[j].x = s[j].y = 0.f;
  }
  VB<QAMD_F64>::st(x, rr, i, h);
  VB<QAMD_F32>::st(xs, rr, i, s);
}

// r64 = b - Ax; r32 = r64; |r64|^2 per active rhs
__global__ void k_bru_resid(Field r64, Field b, Field Ax, Field r32, BatchDev *B, RedScratch rs) {
  const int rr = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (B->active[rr] && i < r64.g->vh) {
    double2 bv[3], av[3];
    float2 sv[3];
    VB<QAMD_F64>::ld(b, rr, i, bv);
    VB<QAMD_F64>::ld(Ax, rr, i, av);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      bv[j].x -= av[j].x;
      bv[j].y -= av[j].y;
      red += bv[j].x * bv[j].x + bv[j].y * bv[j].y;
      sv[j] = make_float2((float)bv[j].x, (float)bv[j].y);
    }
    VB<QAMD_F64>::st(r64, rr, i, bv);
    VB<QAMD_F32>::st(r32, rr, i, sv);
  }
  rhs_partial(red, rs);
}

// ----------------------------------------------- FP64 block CG (baseline)
// x_r = 0, r_r = p_r = b_r, |b_r|^2 partial sums into sets a (slot r) and c
// (slot r), active rhs
__global__ void __launch_bounds__(256) k_bcg64_init(Field x, Field r, Field p, Field b, BatchDev *B, RedScratch ra,
                                                    RedScratch rc) {
  const int rr = blockIdx.y;
  if (!B->active[rr]) return;   // uniform over the block
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < x.g->vh) {
    double2 bv[3], z[3];
    VB<QAMD_F64>::ld(b, rr, i, bv);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      red += bv[j].x * bv[j].x + bv[j].y * bv[j].y;
      z[j] = make_double2(0.0, 0.0);
    }
    VB<QAMD_F64>::st(x, rr, i, z);
    VB<QAMD_F64>::st(r, rr, i, bv);
    VB<QAMD_F64>::st(p, rr, i, bv);
  }
  double a[1] = {red};
  block_sum(a);
  if (threadIdx.x == 0) {
    ra.part[rr * ra.max_blocks + blockIdx.x] = a[0];
    rc.part[rr * rc.max_blocks + blockIdx.x] = a[0];
  }
}

// x_r += alpha_r p_r; r_r -= alpha_r Ap_r; |r_r|^2 partial sums (slot r)
__global__ void __launch_bounds__(256) k_bcg64_xr(Field x, Field r, Field p, Field Ap, BatchDev *B, RedScratch rs) {
  const int rr = blockIdx.y;
  if (!B->active[rr]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < x.g->vh) {
    const double al = B->alpha[rr];
    double2 xv[3], rv[3], pv[3], av[3];
    VB<QAMD_F64>::ld(x, rr, i, xv);
    VB<QAMD_F64>::ld(r, rr, i, rv);
    VB<QAMD_F64>::ld(p, rr, i, pv);
    VB<QAMD_F64>::ld(Ap, rr, i, av);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      xv[j].x = fma(al, pv[j].x, xv[j].x);
      xv[j].y = fma(al, pv[j].y, xv[j].y);
      rv[j].x = fma(-al, av[j].x, rv[j].x);
      rv[j].y = fma(-al, av[j].y, rv[j].y);
      red += rv[j].x * rv[j].x + rv[j].y * rv[j].y;
    }
    VB<QAMD_F64>::st(x, rr, i, xv);
    VB<QAMD_F64>::st(r, rr, i, rv);
  }
  rhs_partial(red, rs);
}

// p_r = r_r + beta_r p_r with |p_r|^2 partial sums (slot r)
__global__ void __launch_bounds__(256) k_bcg64_p(Field p, Field r, BatchDev *B, RedScratch rs) {
  const int rr = blockIdx.y;
  if (!B->active[rr]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < p.g->vh) {
    const double be = B->beta[rr];
    double2 pv[3], rv[3];
    VB<QAMD_F64>::ld(p, rr, i, pv);
    VB<QAMD_F64>::ld(r, rr, i, rv);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      pv[j].x = fma(be, pv[j].x, rv[j].x);
      pv[j].y = fma(be, pv[j].y, rv[j].y);
      red += pv[j].x * pv[j].x + pv[j].y * pv[j].y;
    }
    VB<QAMD_F64>::st(p, rr, i, pv);
  }
  rhs_partial(red, rs);
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct BState {
  BatchDev *B, *Bh;
};

}  // namespace qamd

using namespace qamd;

extern "C" {

// Plain FP64 block CG: textbook CG recurrences for every rhs (the
// cgSolveT iteration, lib/cg.c:43-122, in unscaled form), all vectors FP64,
// the normal operator applied to the whole block by the block Dslash
// (links loaded once for all rhs); each rhs leaves the block when its
// iterated |r|^2 < rsq |b|^2 (checked every 16 iterations).
int qamd_stag_cg_block(const qamd_links *l, qamd_field *x, const qamd_field *b, double mass, double rsq,
                       int maxits, qamd_solve_info *info) {
  if (!l || !x || !b || l->kind != 0 || l->prec != QAMD_F64) {
    set_error("stag_cg_block: needs FP64 doublestored staggered links");
    return -1;
  }
  const int n = b->f.nrhs;
  if (n < 1 || n > 8 || x->f.nrhs != n || b->f.prec != QAMD_F64 || x->f.prec != QAMD_F64 || b->f.ncomp != 3 ||
      x->f.ncomp != 3 || b->f.parity != 0 || x->f.parity != 0 || b->lat != l->lat || x->lat != l->lat) {
    set_error("stag_cg_block: x, b must be even FP64 colour-vector blocks of 1..8 rhs on the links' lattice");
    return -1;
  }
  double t0 = now_s();
  qamd_lattice *lat = l->lat;
  const int nb = (int)((lat->g.vh + 255) / 256);
  const dim3 gb(nb, n);
  const FinOps fops = fin_ops(n);
  qamd_field *r = qamd_field_create(lat, QAMD_F64, 3, n, 0);
  qamd_field *p = qamd_field_create(lat, QAMD_F64, 3, n, 0);
  qamd_field *Ap = qamd_field_create(lat, QAMD_F64, 3, n, 0);
  qamd_field *t = qamd_field_create(lat, QAMD_F64, 3, n, 1);
  BatchDev *B = (BatchDev *)dev_alloc(sizeof(BatchDev)), *Bh = nullptr;
  if (hipHostMalloc((void **)&Bh, sizeof(BatchDev), hipHostMallocDefault) != hipSuccess) Bh = nullptr;
  double m2[MAXB], bb[MAXB];
  int its_done[MAXB] = {0};
  int rc = -1, its = 0;
  do {
    if (!r || !p || !Ap || !t || !B || !Bh || red_reserve(nb)) break;
    if (qamd_field_norm2(b, bb)) break;
    memset(Bh, 0, sizeof(BatchDev));
    Bh->nrhs = n;
    for (int i = 0; i < n; i++) {
      m2[i] = mass * mass;
      Bh->m2[i] = m2[i];
      Bh->active[i] = 1;
    }
    if (hipMemcpyAsync(B, Bh, sizeof(BatchDev), hipMemcpyHostToDevice, stream()) != hipSuccess) break;
    hipLaunchKernelGGL(k_bcg64_init, gb, dim3(256), 0, stream(), x->f, r->f, p->f, b->f, B, red_set(0), red_set(3));
    fops.truer(nb, B, stream());
    if (hipMemcpyAsync(Bh, B, sizeof(BatchDev), hipMemcpyDeviceToHost, stream()) != hipSuccess ||
        hipStreamSynchronize(stream()) != hipSuccess)
      break;
    int nact = 0;
    for (int i = 0; i < n; i++) {
      Bh->active[i] = bb[i] > 0 && Bh->rsq[i] >= rsq * bb[i];
      nact += Bh->active[i];
    }
    if (hipMemcpyAsync(B, Bh, sizeof(BatchDev), hipMemcpyHostToDevice, stream()) != hipSuccess) break;
    int err = 0;
    double zero[MAXB] = {0.0};
    while (nact > 0 && its < maxits && !err) {
      unsigned act = 0;
      for (int i = 0; i < n; i++)
        if (Bh->active[i]) act |= 1u << i;
      for (int k = 0; k < 16 && !err; k++) {
        // t = D_oe p (|t_r|^2), alpha_r = rsq_r / (m2 |p_r|^2 + |t_r|^2),
        // Ap = m2 p - D_eo t, x += alpha p, r -= alpha Ap (|r_r|^2), beta_r,
        // p = r + beta p (|p_r|^2)
        err |= stag_block_launch_m(l, t->f, p->f, zero, nullptr, 1.0, true, red_set(1), stream(), act);
        fops.alpha(nb, B, stream());
        err |= stag_block_launch_m(l, Ap->f, t->f, m2, &p->f, -1.0, false, RedScratch{}, stream(), act);
        hipLaunchKernelGGL(k_bcg64_xr, gb, dim3(256), 0, stream(), x->f, r->f, p->f, Ap->f, B, red_set(0));
        fops.cg(nb, B, stream());
        hipLaunchKernelGGL(k_bcg64_p, gb, dim3(256), 0, stream(), p->f, r->f, B, red_set(3));
        if (hipGetLastError() != hipSuccess) err = 1;
      }
      if (err || hipMemcpyAsync(Bh, B, sizeof(BatchDev), hipMemcpyDeviceToHost, stream()) != hipSuccess ||
          hipStreamSynchronize(stream()) != hipSuccess) {
        err = 1;
        break;
      }
      its += 16;
      nact = 0;
      for (int i = 0; i < n; i++) {
        if (Bh->active[i] && Bh->rsq[i] < rsq * bb[i]) {
          Bh->active[i] = 0;
          its_done[i] = its;
        }
        nact += Bh->active[i];
      }
      if (hipMemcpyAsync(B, Bh, sizeof(BatchDev), hipMemcpyHostToDevice, stream()) != hipSuccess) err = 1;
    }
    if (err) {
      if (!qamd_last_error()[0]) set_error("stag_cg_block: launch failed");
      break;
    }
    if (info)
      for (int i = 0; i < n; i++) {
        memset(&info[i], 0, sizeof info[i]);
        info[i].iterations = Bh->active[i] ? its : its_done[i];
        info[i].iter_rsq = bb[i] > 0 ? Bh->rsq[i] / bb[i] : 0;
        info[i].true_rsq = info[i].iter_rsq;
        info[i].converged = bb[i] == 0 || Bh->rsq[i] < rsq * bb[i];
        info[i].seconds = now_s() - t0;
      }
    rc = 0;
  } while (0);
  qamd_field_destroy(r);
  qamd_field_destroy(p);
  qamd_field_destroy(Ap);
  qamd_field_destroy(t);
  dev_free(B);
  if (Bh) (void)hipHostFree(Bh);
  return rc;
}


int qamd_stag_cg_mixed_block(const qamd_links *l64, const qamd_links *ls_in, qamd_field *x, const qamd_field *b,
                             double mass, double rsq, int maxits, double delta, qamd_solve_info *info) {
  const qamd_links *ls = ls_in;
  if (!l64 || !ls || l64->kind != 0 || ls->kind != 0 || l64->prec != QAMD_F64 || ls->prec == QAMD_F64 ||
      l64->lat->g.V != ls->lat->g.V) {
    set_error("cg_mixed_block: need FP64 links and FP32/FP16/Q16 sloppy links on one lattice");
    return -1;
  }
  const int n = b->f.nrhs;
  if (n < 1 || n > 8 || x->f.nrhs != n || x->f.ncomp != 3 || b->f.ncomp != 3 || x->f.prec != QAMD_F64 ||
      b->f.prec != QAMD_F64 || x->f.parity != 0 || b->f.parity != 0) {
    set_error("cg_mixed_block: x, b must be even FP64 colour-vector blocks of 1..8 rhs");
    return -1;
  }
  // 16-bit links too coarse for this mass: FP32 links for the whole block
  qamd_links *l32 = nullptr;
  if (sloppy_too_coarse(l64, ls, mass, &l32)) return -1;
  double m2r[MAXB];
  for (int i = 0; i < n; i++) m2r[i] = mass * mass;
  int rc = stag_cg_block_core(l64, l32 ? l32 : ls, x, b, m2r, false, rsq, maxits, delta, info);
  if (info && l32)
    for (int i = 0; i < n; i++) info[i].escalations = 1;
  qamd_links_destroy(l32);
  return rc;
}

}  // extern "C"

namespace qamd {

// Block CG on (m2[r] - D_eo D_oe) x_r = b_r, one m2 per rhs, FP32 Krylov
// vectors on sloppy links ls, FP64 reliable updates on l64; x0: start from
// the x passed in (else from zero).  Converged rhs leave the block Dslash
// launches (active mask) as well as the vector kernels.
int stag_cg_block_core(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b,
                       const double *m2, bool x0, double rsq, int maxits, double delta, qamd_solve_info *info) {
  const int n = b->f.nrhs;
  if (delta <= 0 || delta >= 1) delta = 0.1;
  double t0 = now_s();
  qamd_lattice *lat = l64->lat;
  const int ps = ls->prec;
  const bool verbose = getenv("QAMD_VERBOSE") != nullptr;
  const int nb = (int)((lat->g.vh + 255) / 256);
  const dim3 gb(nb, n);
  const FinOps fops = fin_ops(n);
  BState st = {nullptr, nullptr};
  qamd_field *r64 = qamd_field_create(lat, QAMD_F64, 3, n, 0);
  qamd_field *Ax = qamd_field_create(lat, QAMD_F64, 3, n, 0);
  qamd_field *t64 = qamd_field_create(lat, QAMD_F64, 3, n, 1);
  qamd_field *r = qamd_field_create(lat, QAMD_F32, 3, n, 0);
  qamd_field *p = qamd_field_create(lat, QAMD_F32, 3, n, 0);
  qamd_field *xs = qamd_field_create(lat, QAMD_F32, 3, n, 0);
  // D_oe reads the FP32 p directly (16-bit links: the link scale goes onto
  // the link, no 16-bit copy of p); t = D_oe p is stored in the links'
  // format.  32^4, 8 rhs, Q16 links: 0.963 s with a Q16 copy of p, 0.926 s
  // reading FP32 p; an FP32 t as well: 0.98 s.
  qamd_field *t = qamd_field_create(lat, ps, 3, n, 1);
  st.B = (BatchDev *)dev_alloc(sizeof(BatchDev));
  if (hipHostMalloc((void **)&st.Bh, sizeof(BatchDev), hipHostMallocDefault) != hipSuccess) st.Bh = nullptr;
  int rc = -1, nupd = 0, its = 0, nrestart = 0;
  double bb[MAXB], r2[MAXB];
  int its_done[MAXB] = {0};
  auto act_mask = [&]() -> unsigned {
    unsigned m = 0;
    for (int i = 0; i < n; i++)
      if (st.Bh->active[i]) m |= 1u << i;
    return m;
  };
  auto push = [&]() -> int {
    QAMD_HIP(hipMemcpyAsync(st.B, st.Bh, sizeof(BatchDev), hipMemcpyHostToDevice, stream()));
    return 0;
  };
  auto pull = [&]() -> int {
    QAMD_HIP(hipMemcpyAsync(st.Bh, st.B, sizeof(BatchDev), hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    return 0;
  };
  // A v on blocks: t = D v (set 1), out = m2 w - D t (set 2), finisher
  // (the host mirror's active flags are those on the device when this runs)
  auto normal = [&](const qamd_links *L, qamd_field *out, qamd_field *vin, qamd_field *vw, qamd_field *tt) -> int {
    const unsigned act = act_mask();
    double zero[MAXB] = {0.0};
    if (stag_block_launch_m(L, tt->f, vin->f, zero, nullptr, 1.0, true, red_set(1), stream(), act)) return -1;
    if (stag_block_launch_m(L, out->f, tt->f, m2, &vw->f, -1.0, true, red_set(2), stream(), act)) return -1;
    fops.normal(nb, st.B, stream());
    QAMD_HIP(hipGetLastError());
    return 0;
  };
  // p update (mode: 0 p = r + beta p, 1 p = r for the reset rhs, 2 p = r for
  // all active rhs) with |p_r|^2 in partial-sum set 3
  auto p_update = [&](int mode) -> int {
    hipLaunchKernelGGL(k_bcg_pn, gb, dim3(256), 0, stream(), p->f, r->f, st.B, red_set(3), mode);
    QAMD_HIP(hipGetLastError());
    return 0;
  };
  auto reliable = [&]() -> int {
    if (normal(l64, Ax, x, x, t64)) return -1;
    hipLaunchKernelGGL(k_bru_resid, gb, dim3(256), 0, stream(), r64->f, b->f, Ax->f, r->f, st.B, red_set(0));
    fops.truer(nb, st.B, stream());
    QAMD_HIP(hipGetLastError());
    return pull();
  };
  do {
    if (!r64 || !Ax || !t64 || !r || !p || !xs || !t || !st.B || !st.Bh) break;
    if (red_reserve(nb)) break;
    if (qamd_field_norm2(b, bb)) break;
    if (!x0 && qamd_field_zero(x)) break;
    memset(st.Bh, 0, sizeof(BatchDev));
    for (int i = 0; i < n; i++) st.Bh->m2[i] = m2[i];
    st.Bh->nrhs = n;
    for (int i = 0; i < n; i++) st.Bh->active[i] = 1;
    if (push() || reliable()) break;
    int nact = 0;
    for (int i = 0; i < n; i++) {
      r2[i] = st.Bh->rsq[i];
      st.Bh->active[i] = bb[i] > 0 && r2[i] >= rsq * bb[i];
      nact += st.Bh->active[i];
    }
    if (push()) break;
    if (qamd_field_zero(xs) || p_update(2)) break;
    int err = 0;
    while (nact > 0 && its < maxits && !err) {
      // per iteration: t_r = D_oe p_r (|t_r|^2); alpha_r; the fused step
      // x_r += alpha_r p_r, r_r -= alpha_r (m2_r p_r - D_eo t_r), |r_r|^2;
      // beta_r; p_r = r_r + beta_r p_r with |p_r|^2
      const unsigned act = act_mask();
      double zero[MAXB] = {0.0};
      for (int k = 0; k < 16 && !err; k++) {
        err |= stag_block_launch_m(ls, t->f, p->f, zero, nullptr, 1.0, true, red_set(1), stream(), act);
        fops.alpha(nb, st.B, stream());
        err |= stag_block_cg_step(ls, t->f, p->f, xs->f, r->f, m2, st.B->alpha, red_set(0), stream(), act);
        fops.cg(nb, st.B, stream());
        err |= p_update(0);
        if (hipGetLastError() != hipSuccess) err = 1;
      }
      if (err || pull()) {
        err = 1;
        break;
      }
      its += 16;
      // an update when any active rhs dropped by delta, converged, or grew
      // well above its last true residual; all active rhs take it together
      bool upd = its >= maxits;
      bool trig[MAXB];
      for (int i = 0; i < n; i++) {
        double q = st.Bh->rsq[i];
        trig[i] = st.Bh->active[i] &&
                  (!isfinite(q) || sqrt(q) < delta * st.Bh->maxrr[i] || q < rsq * bb[i] || q > 4.0 * r2[i]);
        upd |= trig[i];
      }
      if (verbose)
        fprintf(stderr, "cg_block its %d rel|r0| %.3e\n", its, sqrt(st.Bh->rsq[0] / bb[0]));
      if (!upd) continue;
      double r2s[MAXB];
      for (int i = 0; i < n; i++) r2s[i] = st.Bh->rsq[i];
      hipLaunchKernelGGL(k_bru_accum, gb, dim3(256), 0, stream(), x->f, xs->f, st.B);
      if (hipGetLastError() != hipSuccess || reliable()) {
        err = 1;
        break;
      }
      nupd++;
      // restart p_r = r_r where the true residual left the iterated one
      // behind, or where a delta drop did not carry over to the true residual
      int nreset = 0;
      nact = 0;
      for (int i = 0; i < n; i++) {
        double q = st.Bh->rsq[i];
        if (!isfinite(q)) {
          set_error("cg_mixed_block: rhs %d is not finite after %d iterations", i, its);
          err = 1;
        }
        bool was = st.Bh->active[i];
        if (q < rsq * bb[i]) st.Bh->active[i] = 0;
        if (was && !st.Bh->active[i]) its_done[i] = its;
        st.Bh->reset[i] = st.Bh->active[i] && (q > 2.0 * r2s[i] || (trig[i] && q > 0.25 * r2[i]));
        nreset += st.Bh->reset[i];
        if (was) r2[i] = q;
        nact += st.Bh->active[i];
      }
      if (err) break;
      if (push()) {
        err = 1;
        break;
      }
      if (nreset) {
        if (p_update(1)) err = 1;
        nrestart += nreset;
      }
    }
    if (err) break;
    if (info)
      for (int i = 0; i < n; i++) {
        memset(&info[i], 0, sizeof info[i]);
        info[i].iterations = st.Bh->active[i] ? its : its_done[i];
        info[i].reliable_updates = nupd;
        info[i].true_rsq = bb[i] > 0 ? st.Bh->rsq[i] / bb[i] : 0;
        info[i].iter_rsq = info[i].true_rsq;
        info[i].converged = bb[i] == 0 || st.Bh->rsq[i] < rsq * bb[i];
        info[i].restarts = nrestart;
        info[i].seconds = now_s() - t0;
      }
    rc = 0;
  } while (0);
  qamd_field_destroy(r64);
  qamd_field_destroy(Ax);
  qamd_field_destroy(t64);
  qamd_field_destroy(r);
  qamd_field_destroy(p);
  qamd_field_destroy(xs);
  qamd_field_destroy(t);
  dev_free(st.B);
  if (st.Bh) (void)hipHostFree(st.Bh);
  return rc;
}

}  // namespace qamd
