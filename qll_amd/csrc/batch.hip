// Many independent asqtad solves at once: problems on different lattices,
// with their own gauge fields, masses and numbers of rhs, advance together in
// fused multi-problem kernels.  Every problem owns a contiguous range of
// 256-site blocks in one global numbering; a fused launch maps its block to
// (problem, local block) through a small device table, reads that problem's
// arguments from device memory, and writes its partial sums at the global
// block index; one workgroup per problem then finishes the sums and runs the
// per-rhs CG scalars.  Problems are grouped by rhs count (1, 2, 4, 8) so that
// each group runs a kernel compiled for its register footprint.  The 16
// iterations between two host checks can be captured once in a hipGraph and
// replayed.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"
#include "qamd_vec3.h"
#include "qamd_block.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <chrono>
#include <vector>

namespace qamd {

constexpr int BMAXP = 64;   // problems per batch
constexpr int BR = 8;       // rhs per problem
constexpr int NCLS = 4;     // rhs classes 1, 2, 4, 8
constexpr int BITS = 16;    // iterations per host check

struct ProbDev {
  Geom g;
  int nrhs, nb, gstart;
  double m2;
  double2 *x, *b, *r64, *Ax;   // even parity, FP64, [rhs][3][vh_pad]
  float2 *r, *p, *xs;          // even parity, FP32
  double rsq[BR], rsq_old[BR], t2[BR], p2[BR], beta[BR], maxrr[BR];
  double alpha[BR];            // this iteration's step per rhs (0 for inactive rhs)
  int active[BR], reset[BR];
  int any_active, ru;          // host-set gates: iterate / take a reliable update
};

struct Plan {
  const unsigned char *blk2prob;   // block of this launch -> problem
  const int *start;                // first block of each problem in this launch
  int nblocks;
};

__device__ __forceinline__ bool gate_ok(const ProbDev &D, int gate) { return gate == 0 ? D.any_active : D.ru; }

// fused block Dslash over the problems of one plan; CGX: the fused CG step
// (x_r += alpha_r p_r, r_r -= alpha_r Ap_r, |r_r|^2 in slot r)
template <int P, int PO, int NT, int NR, bool RED, bool CGX = false, int PI = P>
__global__ void __launch_bounds__(256) k_bat_block(const MArgs<P, PO, PI> *args, Plan pl, const ProbDev *pd,
                                                   int gate) {
  int cblk = xcd_block(blockIdx.x, gridDim.x);
  int p = pl.blk2prob[cblk];
  const ProbDev &D = pd[p];
  if (!gate_ok(D, gate)) return;   // uniform over the block
  int local = cblk - pl.start[p];
  int cb = local * 256 + threadIdx.x;
  double red[2 * NR];
#pragma unroll
  for (int i = 0; i < 2 * NR; i++) red[i] = 0.0;
  if (cb < D.g.vh) block_site<P, PO, NT, NR, RED, CGX, PI>(args[p], cb, red);
  if constexpr (CGX) {
    double rr[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) rr[r] = red[2 * r];
    grid_partial_at<NR>(rr, args[p].rs, D.gstart + local);
  } else if (RED) {
    grid_partial_at<2 * NR>(red, args[p].rs, D.gstart + local);
  }
}

// sum of N slots over a problem's blocks [b0, b0 + nb)
template <int N>
__device__ __forceinline__ void range_sum(RedScratch rs, int b0, int nb, double (&w)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) w[i] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    double v[N];
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = rs.part[i * rs.max_blocks + b0 + b];
#pragma unroll
    for (int i = 0; i < N; i++) w[i] += v[i];
  }
  block_sum(w);
}

// alpha_r = rsq_r / (m^2 |p_r|^2 + |t_r|^2) from |t_r|^2 (set a, slot 2r) and
// |p_r|^2 (set b, slot r, from the p update); 0 for inactive rhs, so the fused
// step leaves their x and r as they are
__global__ void __launch_bounds__(256) k_bat_fin_alpha(RedScratch a, RedScratch b, ProbDev *pd) {
  ProbDev &D = pd[blockIdx.x];
  if (!D.any_active) return;
  double wa[2 * BR], wb[BR];
  range_sum(a, D.gstart, D.nb, wa);
  range_sum(b, D.gstart, D.nb, wb);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < BR; r++)
      if (r < D.nrhs) {
        D.t2[r] = wa[2 * r];
        D.p2[r] = wb[r];
        D.alpha[r] = D.active[r] ? D.rsq[r] / (D.m2 * wb[r] + wa[2 * r]) : 0.0;
      }
}

// |r|^2 per rhs -> rsq, beta, running maximum (active rhs)
__global__ void __launch_bounds__(256) k_bat_fin_cg(RedScratch a, ProbDev *pd) {
  ProbDev &D = pd[blockIdx.x];
  if (!D.any_active) return;
  double w[BR];
  range_sum(a, D.gstart, D.nb, w);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < BR; r++)
      if (r < D.nrhs && D.active[r]) {
        D.rsq_old[r] = D.rsq[r];
        D.rsq[r] = w[r];
        D.beta[r] = D.rsq[r] / D.rsq_old[r];
        D.maxrr[r] = fmax(D.maxrr[r], sqrt(D.rsq[r]));
      }
}

// true |r|^2 per rhs after a reliable update
__global__ void __launch_bounds__(256) k_bat_fin_true(RedScratch a, ProbDev *pd) {
  ProbDev &D = pd[blockIdx.x];
  if (!D.ru) return;
  double w[BR];
  range_sum(a, D.gstart, D.nb, w);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < BR; r++)
      if (r < D.nrhs) {
        D.rsq[r] = w[r];
        D.maxrr[r] = sqrt(w[r]);
      }
}

// block of a plain per-site kernel over all problems: (problem, site)
struct SiteOf {
  int p, i;
};
__device__ __forceinline__ SiteOf site_in(const Plan &all) {
  SiteOf s;
  s.p = all.blk2prob[blockIdx.x];
  s.i = (blockIdx.x - all.start[s.p]) * 256 + threadIdx.x;
  return s;
}

template <class T>
__device__ __forceinline__ T *rhs_ptr(T *base, const ProbDev &D, int r) {
  return base + (int64_t)r * 3 * D.g.vh_pad;
}

// p = r + beta p (active rhs) and |p_r|^2 partial sums (slot r; 0 for
// inactive rhs), which the next alpha needs
__global__ void __launch_bounds__(256) k_bat_p(ProbDev *pd, Plan all, RedScratch rs) {
  SiteOf s = site_in(all);
  const ProbDev &D = pd[s.p];
  if (!D.any_active) return;   // uniform over the block
  const int64_t vp = D.g.vh_pad;
  double red[BR];
#pragma unroll
  for (int r = 0; r < BR; r++) red[r] = 0.0;
  if (s.i < D.g.vh) {
#pragma unroll
    for (int r = 0; r < BR; r++) {
      if (r >= D.nrhs || !D.active[r]) continue;
      float be = (float)D.beta[r];
      float2 pv[3], rv[3];
      VecLd<QAMD_F32>::ld(rhs_ptr(D.p, D, r), nullptr, vp, s.i, pv, 1.f);
      VecLd<QAMD_F32>::ld(rhs_ptr(D.r, D, r), nullptr, vp, s.i, rv, 1.f);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        pv[j].x = fmaf(be, pv[j].x, rv[j].x);
        pv[j].y = fmaf(be, pv[j].y, rv[j].y);
        red[r] += (double)pv[j].x * pv[j].x + (double)pv[j].y * pv[j].y;
      }
      VecSt<QAMD_F32>::st(rhs_ptr(D.p, D, r), nullptr, vp, s.i, pv);
    }
  }
  grid_partial_at<BR>(red, rs, blockIdx.x);
}

// restart: p = r for the rhs flagged in reset, with their |p_r|^2 partial
// sums (the other rhs' slots keep the last p update's)
__global__ void __launch_bounds__(256) k_bat_reset(ProbDev *pd, Plan all, RedScratch rs) {
  SiteOf s = site_in(all);
  const ProbDev &D = pd[s.p];
  const int64_t vp = D.g.vh_pad;
  double red[BR];
#pragma unroll
  for (int r = 0; r < BR; r++) red[r] = 0.0;
  if (s.i < D.g.vh) {
#pragma unroll
    for (int r = 0; r < BR; r++) {
      if (r >= D.nrhs || !D.reset[r]) continue;
      float2 rv[3];
      VecLd<QAMD_F32>::ld(rhs_ptr(D.r, D, r), nullptr, vp, s.i, rv, 1.f);
#pragma unroll
      for (int j = 0; j < 3; j++) red[r] += (double)rv[j].x * rv[j].x + (double)rv[j].y * rv[j].y;
      VecSt<QAMD_F32>::st(rhs_ptr(D.p, D, r), nullptr, vp, s.i, rv);
    }
  }
  block_sum(red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < BR; r++)
      if (r < D.nrhs && D.reset[r]) rs.part[r * rs.max_blocks + blockIdx.x] = red[r];
}

// reliable update, first half: x += xs, xs = 0 (active rhs of gated problems)
__global__ void __launch_bounds__(256) k_bat_accum(ProbDev *pd, Plan all) {
  SiteOf s = site_in(all);
  const ProbDev &D = pd[s.p];
  if (!D.ru || s.i >= D.g.vh) return;
  const int64_t vp = D.g.vh_pad;
#pragma unroll
  for (int r = 0; r < BR; r++) {
    if (r >= D.nrhs || !D.active[r]) continue;
    double2 h[3];
    float2 f[3];
    VecLd<QAMD_F64>::ld(rhs_ptr(D.x, D, r), nullptr, vp, s.i, h, 1.0);
    VecLd<QAMD_F32>::ld(rhs_ptr(D.xs, D, r), nullptr, vp, s.i, f, 1.f);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      h[j].x += f[j].x;
      h[j].y += f[j].y;
      f[j].x = f[j].y = 0.f;
    }
    VecSt<QAMD_F64>::st(rhs_ptr(D.x, D, r), nullptr, vp, s.i, h);
    VecSt<QAMD_F32>::st(rhs_ptr(D.xs, D, r), nullptr, vp, s.i, f);
  }
}

// reliable update, second half: r64 = b - Ax, r = r64, |r64|^2 per rhs
__global__ void __launch_bounds__(256) k_bat_resid(ProbDev *pd, Plan all, RedScratch rs) {
  SiteOf s = site_in(all);
  const ProbDev &D = pd[s.p];
  if (!D.ru) return;
  const int64_t vp = D.g.vh_pad;
  double red[BR];
#pragma unroll
  for (int r = 0; r < BR; r++) red[r] = 0.0;
  if (s.i < D.g.vh) {
#pragma unroll
    for (int r = 0; r < BR; r++) {
      if (r >= D.nrhs) continue;
      double2 bv[3], av[3];
      float2 f[3];
      VecLd<QAMD_F64>::ld(rhs_ptr(D.b, D, r), nullptr, vp, s.i, bv, 1.0);
      VecLd<QAMD_F64>::ld(rhs_ptr(D.Ax, D, r), nullptr, vp, s.i, av, 1.0);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        bv[j].x -= av[j].x;
        bv[j].y -= av[j].y;
        red[r] += bv[j].x * bv[j].x + bv[j].y * bv[j].y;
        f[j] = make_float2((float)bv[j].x, (float)bv[j].y);
      }
      VecSt<QAMD_F64>::st(rhs_ptr(D.r64, D, r), nullptr, vp, s.i, bv);
      VecSt<QAMD_F32>::st(rhs_ptr(D.r, D, r), nullptr, vp, s.i, f);
    }
  }
  grid_partial_at<BR>(red, rs, blockIdx.x);
}

static double bnow() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
static T *upload(const std::vector<T> &v) {
  T *d = (T *)dev_alloc(v.size() * sizeof(T));
  if (d && hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream()) != hipSuccess) {
    dev_free(d);
    return nullptr;
  }
  return d;
}

// host-built kernel arguments for one link set / field set
template <int P, int PO, int PI = P>
static MArgs<P, PO, PI> margs(const qamd_links *l, int par, const void *in, const float *in_n, void *out,
                              float *out_n, const void *w, const float *w_n, double a, double b, int nr,
                              RedScratch rs) {
  typedef typename Prec<P>::S S;
  typedef typename Prec<PO>::S SO;
  typedef typename Prec<PI>::S SI;
  MArgs<P, PO, PI> A;
  memset(&A, 0, sizeof A);
  A.g = l->lat->g;
  A.links = (const S *)l->data[par];
  A.in = (const SI *)in;
  A.in_norm = in_n;
  A.out = (SO *)out;
  A.out_norm = out_n;
  A.w = (const SO *)w;
  A.w_norm = w_n;
  for (int r = 0; r < 16; r++) {
    A.ar[r] = a;
    A.rmap[r] = (unsigned char)r;
  }
  A.b = b;
  A.par = par;
  A.nr = nr;
  A.rs = rs;
  A.lsign = (l->recon && l->nterms == 16) ? l->lsign[par] : nullptr;
  for (int t = 0; t < 16; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  return A;
}

struct Work {   // per problem, library-owned work fields
  qamd_field *r64, *Ax, *t64, *r, *p, *xs, *t;
  qamd_links *l32;   // FP32 links when the 16-bit ones are too coarse for its mass
};

// one batch solve with sloppy links of precision PS (16 terms)
template <int PS>
struct Runner {
  static constexpr bool HALF = (PS == QAMD_F16 || PS == QAMD_Q16);
  // t = D_oe p reads the FP32 p itself (16-bit links: the term scale goes
  // onto the link)
  static constexpr int PI1 = HALF ? QAMD_F32 : PS;
  MArgs<PS, PS, PI1> *s1 = nullptr;     // group 0: the common sloppy links
  MArgs<PS, QAMD_F32> *s2 = nullptr;
  MArgs<QAMD_F32, QAMD_F32> *f1 = nullptr, *f2 = nullptr;   // group 1: FP32 fallback
  MArgs<QAMD_F64, QAMD_F64> *r1 = nullptr, *r2 = nullptr;
  Plan cls[2][NCLS], all;                // [group][rhs class]
  ProbDev *pd = nullptr;
  int nprob = 0;

  template <int P, int PO, int NR, bool RED, bool CGX, int PI>
  static void bl(const MArgs<P, PO, PI> *a, const Plan &pl, const ProbDev *pd, int gate, hipStream_t st) {
    if (pl.nblocks)
      hipLaunchKernelGGL((k_bat_block<P, PO, 16, NR, RED, CGX, PI>), dim3(pl.nblocks), dim3(256), 0, st, a, pl,
                         pd, gate);
  }
  template <int P, int PO, bool RED, bool CGX = false, int PI = P>
  void bl_all(const MArgs<P, PO, PI> *a, int grp, int gate, hipStream_t st) {
    bl<P, PO, 1, RED, CGX, PI>(a, cls[grp][0], pd, gate, st);
    bl<P, PO, 2, RED, CGX, PI>(a, cls[grp][1], pd, gate, st);
    bl<P, PO, 4, RED, CGX, PI>(a, cls[grp][2], pd, gate, st);
    bl<P, PO, 8, RED, CGX, PI>(a, cls[grp][3], pd, gate, st);
  }
  // t = D_oe p (|t|^2); alpha; the fused step x += alpha p, r -= alpha
  // (m^2 p - D_eo t), |r|^2; beta; p = r + beta p with |p|^2
  void iterate(hipStream_t st) {
    bl_all<PS, PS, true, false, PI1>(s1, 0, 0, st);
    if (PS != QAMD_F32) bl_all<QAMD_F32, QAMD_F32, true>(f1, 1, 0, st);
    hipLaunchKernelGGL(k_bat_fin_alpha, dim3(nprob), dim3(256), 0, st, red_set(1), red_set(3), pd);
    bl_all<PS, QAMD_F32, true, true>(s2, 0, 0, st);
    if (PS != QAMD_F32) bl_all<QAMD_F32, QAMD_F32, true, true>(f2, 1, 0, st);
    hipLaunchKernelGGL(k_bat_fin_cg, dim3(nprob), dim3(256), 0, st, red_set(0), pd);
    hipLaunchKernelGGL(k_bat_p, dim3(all.nblocks), dim3(256), 0, st, pd, all, red_set(3));
  }
  void reliable(hipStream_t st) {
    hipLaunchKernelGGL(k_bat_accum, dim3(all.nblocks), dim3(256), 0, st, pd, all);
    for (int g = 0; g < 2; g++) bl_all<QAMD_F64, QAMD_F64, false>(r1, g, 1, st);
    for (int g = 0; g < 2; g++) bl_all<QAMD_F64, QAMD_F64, false>(r2, g, 1, st);
    hipLaunchKernelGGL(k_bat_resid, dim3(all.nblocks), dim3(256), 0, st, pd, all, red_set(0));
    hipLaunchKernelGGL(k_bat_fin_true, dim3(nprob), dim3(256), 0, st, red_set(0), pd);
  }
  void reset(hipStream_t st) {
    hipLaunchKernelGGL(k_bat_reset, dim3(all.nblocks), dim3(256), 0, st, pd, all, red_set(3));
  }
  void release() {
    dev_free(s1);
    dev_free(s2);
    dev_free(f1);
    dev_free(f2);
    dev_free(r1);
    dev_free(r2);
  }

  int run(const qamd_stag_problem *pr, int n, std::vector<ProbDev> &host, std::vector<Work> &wk, double rsq,
          int maxits, double delta, int use_graph, qamd_solve_info *info);
};

template <int PS>
int Runner<PS>::run(const qamd_stag_problem *pr, int n, std::vector<ProbDev> &H, std::vector<Work> &wk, double rsq,
                    int maxits, double delta, int use_graph, qamd_solve_info *info) {
  const double t0 = bnow();
  const bool verbose = getenv("QAMD_VERBOSE") != nullptr;
  nprob = n;
  // kernel arguments (the scratch was reserved before, so its sets are final)
  std::vector<MArgs<PS, PS, PI1>> a1(n);
  std::vector<MArgs<PS, QAMD_F32>> a2(n);
  std::vector<MArgs<QAMD_F32, QAMD_F32>> c1(n), c2(n);
  std::vector<MArgs<QAMD_F64, QAMD_F64>> b1(n), b2(n);
  for (int i = 0; i < n; i++) {
    const Work &w = wk[i];
    const int nr = H[i].nrhs;
    const double m2 = H[i].m2;
    if (!w.l32) {
      a1[i] = margs<PS, PS, PI1>(pr[i].ls, 1, w.p->f.data, nullptr, w.t->f.data, w.t->f.norm, nullptr, nullptr,
                                 0.0, 1.0, nr, red_set(1));
      a2[i] = margs<PS, QAMD_F32>(pr[i].ls, 0, w.t->f.data, w.t->f.norm, nullptr, nullptr, w.p->f.data,
                                  nullptr, m2, -1.0, nr, red_set(0));
      a2[i].cgx = (float2 *)w.xs->f.data;
      a2[i].cgr = (float2 *)w.r->f.data;
      a2[i].alpha = pd[i].alpha;
    } else {
      c1[i] = margs<QAMD_F32, QAMD_F32>(w.l32, 1, w.p->f.data, nullptr, w.t->f.data, nullptr, nullptr, nullptr, 0.0,
                                        1.0, nr, red_set(1));
      c2[i] = margs<QAMD_F32, QAMD_F32>(w.l32, 0, w.t->f.data, nullptr, nullptr, nullptr, w.p->f.data, nullptr,
                                        m2, -1.0, nr, red_set(0));
      c2[i].cgx = (float2 *)w.xs->f.data;
      c2[i].cgr = (float2 *)w.r->f.data;
      c2[i].alpha = pd[i].alpha;
    }
    b1[i] = margs<QAMD_F64, QAMD_F64>(pr[i].l64, 1, pr[i].x->f.data, nullptr, w.t64->f.data, nullptr, nullptr,
                                      nullptr, 0.0, 1.0, nr, red_set(3));
    b2[i] = margs<QAMD_F64, QAMD_F64>(pr[i].l64, 0, w.t64->f.data, nullptr, w.Ax->f.data, nullptr,
                                      pr[i].x->f.data, nullptr, m2, -1.0, nr, red_set(3));
  }
  s1 = upload(a1);
  s2 = upload(a2);
  f1 = upload(c1);
  f2 = upload(c2);
  r1 = upload(b1);
  r2 = upload(b2);
  if (!s1 || !s2 || !f1 || !f2 || !r1 || !r2) return -1;

  ProbDev *hp = nullptr;   // pinned mirror of the device state
  if (hipHostMalloc((void **)&hp, n * sizeof(ProbDev), hipHostMallocDefault) != hipSuccess) {
    set_error("cg_mixed_batch: pinned allocation failed");
    return -1;
  }
  memcpy(hp, H.data(), n * sizeof(ProbDev));
  // the copy from pinned memory runs when the stream reaches it, so wait for
  // it before the host writes the mirror again
  auto push = [&]() -> int {
    QAMD_HIP(hipMemcpyAsync(pd, hp, n * sizeof(ProbDev), hipMemcpyHostToDevice, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    return 0;
  };
  auto pull = [&]() -> int {
    QAMD_HIP(hipMemcpyAsync(hp, pd, n * sizeof(ProbDev), hipMemcpyDeviceToHost, stream()));
    QAMD_HIP(hipStreamSynchronize(stream()));
    return 0;
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Capture BITS iterations into a graph.  On any failure the stream is
  // taken out of capture mode, partial graph objects are destroyed and false
  // is returned, so the solve continues with eager launches and every later
  // library call runs normally.  QAMD_TEST_FAIL_CAPTURE=1 makes instantiation
  // report a failure (host-side; tests the recovery without a GPU fault).
  auto capture = [&](hipGraph_t &g, hipGraphExec_t &x) -> bool {
    g = nullptr;
    x = nullptr;
    hipError_t e = hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int k = 0; k < BITS; k++) iterate(stream());
      e = hipStreamEndCapture(stream(), &g);
      const char *inj = getenv("QAMD_TEST_FAIL_CAPTURE");
      if (e == hipSuccess && inj && inj[0] == '1') e = hipErrorInvalidValue;
      if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    }
    if (e == hipSuccess) return true;
    set_error("cg_mixed_batch: graph capture: %s", hipGetErrorString(e));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream(), &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
      hipGraph_t g2 = nullptr;
      (void)hipStreamEndCapture(stream(), &g2);
      if (g2) (void)hipGraphDestroy(g2);
    }
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
    (void)hipGetLastError();
    return false;
  };
  int rc = -1;
  std::vector<int> its(n, 0), nupd(n, 0), nrst(n, 0);
  std::vector<double> bb(n * BR, 0.0), r2(n * BR, 0.0);
  do {
    // x = 0, r = b, |b|^2 through one reliable update of every problem
    for (int i = 0; i < n; i++) {
      hp[i].any_active = 0;
      hp[i].ru = 1;
      for (int r = 0; r < BR; r++) hp[i].active[r] = hp[i].reset[r] = 0;
    }
    if (push()) break;
    reliable(stream());
    if (hipGetLastError() != hipSuccess || pull()) break;
    int nact = 0;
    for (int i = 0; i < n; i++) {
      hp[i].ru = 0;
      hp[i].any_active = 0;
      for (int r = 0; r < hp[i].nrhs; r++) {
        bb[i * BR + r] = r2[i * BR + r] = hp[i].rsq[r];
        hp[i].active[r] = hp[i].reset[r] = bb[i * BR + r] > 0;   // x = 0 is exact for b = 0
        hp[i].any_active |= hp[i].active[r];
      }
      nact += hp[i].any_active;
    }
    if (push()) break;
    reset(stream());
    for (int i = 0; i < n; i++)
      for (int r = 0; r < BR; r++) hp[i].reset[r] = 0;
    if (push()) break;
    if (use_graph && nact && !capture(graph, exec) && verbose)
      fprintf(stderr, "cg_batch: graph capture failed (%s); eager launches\n", qamd_last_error());
    int err = 0;
    while (nact > 0 && !err) {
      if (exec) {
        if (hipGraphLaunch(exec, stream()) != hipSuccess) err = 1;
      } else {
        for (int k = 0; k < BITS; k++) iterate(stream());
        if (hipGetLastError() != hipSuccess) err = 1;
      }
      if (err || pull()) {
        err = 1;
        break;
      }
      if (verbose) {
        int na = 0, nr = 0;
        long sites = 0;
        for (int i = 0; i < n; i++)
          if (hp[i].any_active) {
            na++;
            sites += (long)hp[i].g.vh;
            for (int r = 0; r < hp[i].nrhs; r++) nr += hp[i].active[r];
          }
        fprintf(stderr, "cg_batch %.3f s: %d problems, %d rhs, %ld even sites active\n", bnow() - t0, na, nr, sites);
      }
      // which problems take a reliable update
      int nru = 0;
      std::vector<unsigned char> trig(n * BR, 0);
      for (int i = 0; i < n; i++) {
        if (!hp[i].any_active) continue;
        its[i] += BITS;
        bool upd = its[i] >= maxits;
        for (int r = 0; r < hp[i].nrhs; r++) {
          if (!hp[i].active[r]) continue;
          double q = hp[i].rsq[r];
          bool t = !isfinite(q) || sqrt(q) < delta * hp[i].maxrr[r] || q < rsq * bb[i * BR + r] ||
                   q > 4.0 * r2[i * BR + r];
          trig[i * BR + r] = t;
          upd |= t;
        }
        hp[i].ru = upd;
        nru += upd;
      }
      if (!nru) continue;
      std::vector<double> r2s(n * BR);
      for (int i = 0; i < n; i++)
        for (int r = 0; r < BR; r++) r2s[i * BR + r] = hp[i].rsq[r];
      if (push()) {
        err = 1;
        break;
      }
      reliable(stream());
      if (hipGetLastError() != hipSuccess || pull()) {
        err = 1;
        break;
      }
      int nreset = 0;
      nact = 0;
      for (int i = 0; i < n; i++) {
        if (hp[i].ru) {
          nupd[i]++;
          for (int r = 0; r < hp[i].nrhs; r++) {
            double q = hp[i].rsq[r];
            if (!isfinite(q)) {
              set_error("cg_mixed_batch: problem %d rhs %d is not finite after %d iterations", i, r, its[i]);
              err = 1;
            }
            bool was = hp[i].active[r];
            if (q < rsq * bb[i * BR + r]) hp[i].active[r] = 0;
            hp[i].reset[r] = hp[i].active[r] &&
                             (q > 2.0 * r2s[i * BR + r] || (trig[i * BR + r] && q > 0.25 * r2[i * BR + r]));
            nreset += hp[i].reset[r];
            nrst[i] += hp[i].reset[r];
            if (was) r2[i * BR + r] = q;
          }
          if (its[i] >= maxits)
            for (int r = 0; r < BR; r++) hp[i].active[r] = hp[i].reset[r] = 0;
          hp[i].any_active = 0;
          for (int r = 0; r < hp[i].nrhs; r++) hp[i].any_active |= hp[i].active[r];
          hp[i].ru = 0;
        }
        nact += hp[i].any_active;
      }
      if (err || push()) {
        err = 1;
        break;
      }
      if (nreset) reset(stream());
      for (int i = 0; i < n; i++)
        for (int r = 0; r < BR; r++) hp[i].reset[r] = 0;
      if (nreset && push()) {
        err = 1;
        break;
      }
    }
    if (err) break;
    if (info) {
      double secs = bnow() - t0;
      int k = 0;
      for (int i = 0; i < n; i++)
        for (int r = 0; r < hp[i].nrhs; r++, k++) {
          memset(&info[k], 0, sizeof info[k]);
          double b2 = bb[i * BR + r];
          info[k].iterations = its[i];
          info[k].reliable_updates = nupd[i];
          info[k].restarts = nrst[i];
          info[k].true_rsq = b2 > 0 ? hp[i].rsq[r] / b2 : 0.0;
          info[k].iter_rsq = info[k].true_rsq;
          info[k].converged = b2 == 0 || hp[i].rsq[r] < rsq * b2;
          info[k].escalations = wk[i].l32 != nullptr;
          info[k].seconds = secs;
        }
    }
    rc = 0;
  } while (0);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipHostFree(hp);
  return rc;
}

}  // namespace qamd

using namespace qamd;

extern "C" {

int qamd_stag_cg_mixed_batch(const qamd_stag_problem *pr, int n, double rsq, int maxits, double delta,
                             int use_graph, qamd_solve_info *info) {
  if (!pr || n < 1 || n > BMAXP) {
    set_error("cg_mixed_batch: 1..%d problems", BMAXP);
    return -1;
  }
  const int ps = pr[0].ls ? pr[0].ls->prec : -1;
  for (int i = 0; i < n; i++) {
    const qamd_stag_problem &q = pr[i];
    if (!q.l64 || !q.ls || !q.x || !q.b || q.l64->kind != 0 || q.ls->kind != 0 || q.l64->prec != QAMD_F64 ||
        q.ls->prec != ps || ps == QAMD_F64 || q.l64->nterms != 16 || q.ls->nterms != 16 ||
        q.l64->lat->g.V != q.ls->lat->g.V) {
      set_error("cg_mixed_batch: problem %d needs 16-term FP64 links and sloppy links of one common precision", i);
      return -1;
    }
    const Field &x = q.x->f, &b = q.b->f;
    if (x.prec != QAMD_F64 || b.prec != QAMD_F64 || x.ncomp != 3 || b.ncomp != 3 || x.parity != 0 ||
        b.parity != 0 || x.nrhs != b.nrhs || x.nrhs < 1 || x.nrhs > BR || q.x->lat->g.V != q.l64->lat->g.V ||
        q.b->lat->g.V != q.l64->lat->g.V) {
      set_error("cg_mixed_batch: problem %d: x, b must be even FP64 colour vectors of 1..%d rhs on its lattice", i, BR);
      return -1;
    }
  }
  if (delta <= 0 || delta >= 1) delta = 0.1;
  std::vector<ProbDev> H(n);
  std::vector<Work> wk(n);
  memset(H.data(), 0, n * sizeof(ProbDev));
  memset(wk.data(), 0, n * sizeof(Work));
  auto cleanup = [&]() {
    for (int i = 0; i < n; i++) {
      Work &w = wk[i];
      qamd_field_destroy(w.r64);
      qamd_field_destroy(w.Ax);
      qamd_field_destroy(w.t64);
      qamd_field_destroy(w.r);
      qamd_field_destroy(w.p);
      qamd_field_destroy(w.xs);
      qamd_field_destroy(w.t);
      qamd_links_destroy(w.l32);
    }
  };
  // group 1: problems whose mass is too light for the 16-bit links
  for (int i = 0; i < n; i++)
    if (sloppy_too_coarse(pr[i].l64, pr[i].ls, pr[i].mass, &wk[i].l32)) {
      cleanup();
      return -1;
    }
  // blocks: global numbering (problem order) and one launch plan per
  // (precision group, rhs class)
  constexpr int NPL = 2 * NCLS;
  int total = 0;
  std::vector<int> pl_of(n), cstart(NPL * BMAXP, 0), gstart(BMAXP, 0), cblocks(NPL, 0);
  for (int i = 0; i < n; i++) {
    const Geom &g = pr[i].l64->lat->g;
    H[i].g = g;
    H[i].nrhs = pr[i].x->f.nrhs;
    H[i].nb = (int)((g.vh + 255) / 256);
    H[i].gstart = gstart[i] = total;
    H[i].m2 = pr[i].mass * pr[i].mass;
    total += H[i].nb;
    int c = H[i].nrhs <= 1 ? 0 : H[i].nrhs <= 2 ? 1 : H[i].nrhs <= 4 ? 2 : 3;
    int k = (wk[i].l32 ? NCLS : 0) + c;
    pl_of[i] = k;
    cstart[k * BMAXP + i] = cblocks[k];
    cblocks[k] += H[i].nb;
  }
  std::vector<unsigned char> b2p_all(total);
  std::vector<std::vector<unsigned char>> b2p(NPL);
  for (int k = 0; k < NPL; k++) b2p[k].resize(cblocks[k] > 0 ? cblocks[k] : 1);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < H[i].nb; k++) {
      b2p_all[H[i].gstart + k] = (unsigned char)i;
      b2p[pl_of[i]][cstart[pl_of[i] * BMAXP + i] + k] = (unsigned char)i;
    }
  }
  if (red_reserve(total)) {
    cleanup();
    return -1;
  }
  int rc = -1;
  ProbDev *pd = nullptr;
  unsigned char *d_all = nullptr, *d_cls[NPL] = {nullptr};
  int *d_gstart = nullptr, *d_cstart = nullptr;
  do {
    bool ok = true;
    for (int i = 0; i < n && ok; i++) {
      qamd_lattice *lat = pr[i].l64->lat;
      int nr = H[i].nrhs;
      Work &w = wk[i];
      const int pt = w.l32 ? QAMD_F32 : ps;   // precision of the odd temporary
      w.r64 = qamd_field_create(lat, QAMD_F64, 3, nr, 0);
      w.Ax = qamd_field_create(lat, QAMD_F64, 3, nr, 0);
      w.t64 = qamd_field_create(lat, QAMD_F64, 3, nr, 1);
      w.r = qamd_field_create(lat, QAMD_F32, 3, nr, 0);
      w.p = qamd_field_create(lat, QAMD_F32, 3, nr, 0);
      w.xs = qamd_field_create(lat, QAMD_F32, 3, nr, 0);
      w.t = qamd_field_create(lat, pt, 3, nr, 1);
      ok = w.r64 && w.Ax && w.t64 && w.r && w.p && w.xs && w.t;
      ok = ok && !qamd_field_zero(pr[i].x) && !qamd_field_zero(w.xs) && !qamd_field_zero(w.p);
      if (!ok) break;
      ProbDev &D = H[i];
      D.x = (double2 *)pr[i].x->f.data;
      D.b = (double2 *)pr[i].b->f.data;
      D.r64 = (double2 *)w.r64->f.data;
      D.Ax = (double2 *)w.Ax->f.data;
      D.r = (float2 *)w.r->f.data;
      D.p = (float2 *)w.p->f.data;
      D.xs = (float2 *)w.xs->f.data;
    }
    if (!ok) break;
    pd = upload(H);
    d_all = upload(b2p_all);
    d_gstart = upload(gstart);
    d_cstart = upload(cstart);
    for (int k = 0; k < NPL; k++) ok = ok && (d_cls[k] = upload(b2p[k])) != nullptr;
    if (!pd || !d_all || !d_gstart || !d_cstart || !ok) break;
    auto go = [&](auto &R) -> int {
      R.pd = pd;
      R.all = Plan{d_all, d_gstart, total};
      for (int k = 0; k < NPL; k++) R.cls[k / NCLS][k % NCLS] = Plan{d_cls[k], d_cstart + k * BMAXP, cblocks[k]};
      int r = R.run(pr, n, H, wk, rsq, maxits, delta, use_graph, info);
      R.release();
      return r;
    };
    if (ps == QAMD_F32) {
      Runner<QAMD_F32> R;
      rc = go(R);
    } else if (ps == QAMD_F16) {
      Runner<QAMD_F16> R;
      rc = go(R);
    } else {
      Runner<QAMD_Q16> R;
      rc = go(R);
    }
  } while (0);
  if (hipStreamSynchronize(stream()) != hipSuccess && rc == 0) {
    set_error("cg_mixed_batch: stream error");
    rc = -1;
  }
  dev_free(pd);
  dev_free(d_all);
  dev_free(d_gstart);
  dev_free(d_cstart);
  for (int k = 0; k < NPL; k++) dev_free(d_cls[k]);
  cleanup();
  return rc;
}

}  // extern "C"
