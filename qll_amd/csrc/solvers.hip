// Krylov solvers on the device: QLL's CG and multishift CG recurrences, and
// mixed-precision CG with reliable updates.
//
// Per CG iteration the work is two Dslash launches (the even-odd normal
// operator, with |Dp|^2 and |p|^2 reduced in their epilogues) and two fused
// vector kernels.  All scalars live in a device-side state block: the
// reductions' last block updates them, the next kernels read them, so the host
// never sits inside an iteration; it enqueues iterations in batches and reads
// the state once per batch.  Every kernel of the plain solvers checks the
// state's 'done' flag and returns at once when it is set, which stops the
// batch exactly where the reference's loop stops.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"
#include "qamd_vec3.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <chrono>

namespace qamd {

constexpr int MAXSHIFT = 32;
constexpr int BATCH = 16;   // iterations enqueued between two reads of the state

struct SolverDev {
  double rsq, rsq_old;     // current and previous |r|^2
  double nrm[4];           // Dslash reductions: |t|^2, -, |Ap|^2, |p|^2
  double m2;               // mass^2 of the normal operator
  double rsqstop;
  double alpha, beta;
  double maxrr;            // max |r| since the last reliable update
  int its, itnlim, done, mode;   // mode 0: QLL rescaled CG, 1: textbook CG
  int auto_stop;
  // multishift scalars (cgmsSolveT, lib/cgms.c:6-179)
  int nshift;
  double shifts[MAXSHIFT];
  double a[MAXSHIFT], b[MAXSHIFT], s[MAXSHIFT], bo[MAXSHIFT];
  double z[MAXSHIFT], zo[MAXSHIFT], zn[MAXSHIFT];
  double pa[MAXSHIFT], pzn[MAXSHIFT];   // coefficients for the p update
  double xb[MAXSHIFT];                  // coefficients for the x update
  int xrun;                             // 1: this iteration's x update is due
  int drop;                             // 1: shifts stop updating once converged
  CGScal cg;   // single-reduction CG (mixed solver)
};

template <int P> struct VT {
  typedef typename Prec<P>::S S;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  __device__ static void ld(const Field &f, int i, C (&v)[3]) {
    VecLd<P>::ld((const S *)f.data, f.norm, f.vp, i, v, (R)1);
  }
  __device__ static void st(Field &f, int i, const C (&v)[3]) {
    VecSt<P>::st((S *)f.data, f.norm, f.vp, i, v);
  }
};

// ---------------------------------------------------------------- CG kernels
// p update of the plain CG
template <int P>
__global__ void __launch_bounds__(256) k_cg_p(Field p, Field r, SolverDev *S) {
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  if (S->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.g->vh) return;
  C pv[3], rv[3];
  VT<P>::ld(p, i, pv);
  VT<P>::ld(r, i, rv);
  if (S->mode == 0) {   // p += r / rsq  (lib/cg.c:86-87)
    R be = (R)(1.0 / S->rsq);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      pv[j].x = fma(be, rv[j].x, pv[j].x);
      pv[j].y = fma(be, rv[j].y, pv[j].y);
    }
  } else {              // p = r + beta p
    R be = (R)S->beta;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      pv[j].x = fma(be, pv[j].x, rv[j].x);
      pv[j].y = fma(be, pv[j].y, rv[j].y);
    }
  }
  VT<P>::st(p, i, pv);
}

// Single-reduction CG (Chronopoulos & Gear, J. Comput. Appl. Math. 25
// (1989) 153-168) on A = m^2 - D_eo D_oe: from gamma = |r|^2 (set a) and
// |t|^2 = |D_oe r|^2 (set b), delta = r.Ar = m^2 gamma + |t|^2 (the dslashp2
// identity, lib/stagDslash.c:546-575), beta = gamma / gamma_old and
// alpha = gamma / (delta - beta gamma / alpha_old); a restart takes beta = 0,
// alpha = gamma / delta.
struct FinCG {
  SolverDev *S;
  __device__ void operator()(const double *t) const {
    S->rsq_old = S->rsq;
    S->rsq = t[0];
    S->beta = S->rsq / S->rsq_old;
    S->maxrr = fmax(S->maxrr, sqrt(S->rsq));
    S->its += 1;
    if (S->its >= S->itnlim || (S->auto_stop && S->rsq < S->rsqstop)) S->done = 1;
  }
};

// x += alpha p; r -= alpha Ap; |r|^2   (lib/cg.c:96-110)
template <int P, int PX>
__global__ void __launch_bounds__(256) k_cg_xr(Field x, Field p, Field r, Field Ap, SolverDev *S, RedScratch rs) {
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  typedef typename Prec<PX>::C CX;
  typedef typename Prec<PX>::R RX;
  if (S->done) return;
  double pAp = S->m2 * S->nrm[3] + S->nrm[0];
  double alpha = (S->mode == 0) ? 1.0 / pAp : S->rsq / pAp;
  if (blockIdx.x == 0 && threadIdx.x == 0) S->alpha = alpha;   // read only by the host
  double red[1] = {0.0};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.g->vh) {
    C pv[3], rv[3], av[3];
    CX xv[3];
    VT<P>::ld(p, i, pv);
    VT<P>::ld(r, i, rv);
    VT<P>::ld(Ap, i, av);
    VT<PX>::ld(x, i, xv);
    R al = (R)alpha;
    RX alx = (RX)alpha;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      xv[j].x = fma(alx, (RX)pv[j].x, xv[j].x);
      xv[j].y = fma(alx, (RX)pv[j].y, xv[j].y);
      rv[j].x = fma(-al, av[j].x, rv[j].x);
      rv[j].y = fma(-al, av[j].y, rv[j].y);
    }
    VT<PX>::st(x, i, xv);
    VT<P>::st(r, i, rv);
    // the residual norm is taken of the stored (rounded) vector
    C rr[3];
    VT<P>::ld(r, i, rr);
#pragma unroll
    for (int j = 0; j < 3; j++) red[0] += (double)rr[j].x * rr[j].x + (double)rr[j].y * rr[j].y;
  }
  grid_partial<1>(red, rs);
}

// reliable update, part 1: xhi += xs; xs = 0
template <int PX>
__global__ void k_ru_accum(Field xhi, Field xs) {
  typedef typename Prec<PX>::C CX;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= xhi.g->vh) return;
  double2 h[3];
  CX s[3];
  VT<QAMD_F64>::ld(xhi, i, h);
  VT<PX>::ld(xs, i, s);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    h[j].x += (double)s[j].x;
    h[j].y += (double)s[j].y;
    s[j].x = 0;
    s[j].y = 0;
  }
  VT<QAMD_F64>::st(xhi, i, h);
  VT<PX>::st(xs, i, s);
}

struct FinTrue {
  SolverDev *S;
  __device__ void operator()(const double *t) const {
    S->rsq = t[0];
    S->beta = S->rsq / S->rsq_old;
    S->maxrr = sqrt(S->rsq);
  }
};

// r.s, p.s and r.p (real parts) of FP32 fields, for a reliable update that
// keeps p
__global__ void __launch_bounds__(256) k_cg_keep_dots(Field r, Field p, Field s, RedScratch rs) {
  double red[3] = {0.0, 0.0, 0.0};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < r.g->vh) {
    float2 rv[3], pv[3], sv[3];
    VT<QAMD_F32>::ld(r, i, rv);
    VT<QAMD_F32>::ld(p, i, pv);
    VT<QAMD_F32>::ld(s, i, sv);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      red[0] += (double)rv[j].x * sv[j].x + (double)rv[j].y * sv[j].y;
      red[1] += (double)pv[j].x * sv[j].x + (double)pv[j].y * sv[j].y;
      red[2] += (double)rv[j].x * pv[j].x + (double)rv[j].y * pv[j].y;
    }
  }
  grid_partial<3>(red, rs);
}
struct FinKeep {
  SolverDev *S;
  int ph;
  __device__ void operator()(const double *t) const {
    S->cg.kr[ph] = t[0];
    S->cg.kp[ph] = t[1];
    S->cg.kq[ph] = t[2];
    S->cg.keep[ph] = 1;
    S->cg.maxrr = sqrt(S->rsq);
  }
};

// reliable update, part 2: rhi = b - Ax; rs = rhi; |rhi|^2
template <int PS>
__global__ void k_ru_resid(Field rhi, Field b, Field Ax, Field rs, SolverDev *S, RedScratch red_s) {
  typedef typename Prec<PS>::C CS;
  typedef typename Prec<PS>::R RS;
  double red[1] = {0.0};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rhi.g->vh) {
    double2 bv[3], av[3];
    VT<QAMD_F64>::ld(b, i, bv);
    VT<QAMD_F64>::ld(Ax, i, av);
    CS sv[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      bv[j].x -= av[j].x;
      bv[j].y -= av[j].y;
      red[0] += bv[j].x * bv[j].x + bv[j].y * bv[j].y;
      sv[j].x = (RS)bv[j].x;
      sv[j].y = (RS)bv[j].y;
    }
    VT<QAMD_F64>::st(rhi, i, bv);
    VT<PS>::st(rs, i, sv);
  }
  grid_partial<1>(red, red_s);
}

// ------------------------------------------------------- multishift kernels
// coefficients of cgmsSolveT (lib/cgms.c:78-89) from the current state
// shift i >= 1 of the multishift recurrences (lib/cgms.c:79-88): the new
// zeta_i and beta_i from beta_0 of this iteration and the previous iteration's
// alpha_0, beta_0 (a00, bo0), computed per shift so nothing is indexed at run time
__device__ inline void ms_coef(const SolverDev *S, int i, double b0, double a00, double bo0, double &bi,
                               double &zni) {
  double z = S->z[i], zo = S->zo[i];
  zni = z * zo * bo0;
  double c1 = b0 * a00 * (zo - z);
  c1 += zo * bo0 * (1 + S->shifts[i] * b0);
  if (c1 != 0.0) zni /= c1;
  else zni = 0;
  if (z != 0.0) bi = b0 * zni / z;
  else zni = bi = 0;
}

struct MSPtrs {
  void *x[MAXSHIFT];
  float *xn[MAXSHIFT];
  void *p[MAXSHIFT];
  float *pn[MAXSHIFT];
};

// After the |r|^2 reduction: the iteration's x-update coefficients b_i s_i
// (the x_i update of lib/cgms.c:91-96, applied by k_ms_xp together with the
// p update), the stopping test, and the p-update coefficients.  With drop
// set, a shift i > 0 whose residual |zeta_i r| is below the stopping
// residual takes no further updates (no counterpart in QLL, which iterates
// every shift to the end).
struct FinMS {
  SolverDev *S;
  __device__ void operator()(const double *t) const {
    double pkp = (S->m2 * S->nrm[3] + S->nrm[0]) * S->s[0] * S->s[0];
    const double b0 = S->rsq / pkp, a00 = S->a[0], bo0 = S->bo[0];
    double oldrsq = S->rsq;
    S->rsq = t[0];
    S->its += 1;
    if (S->its >= S->itnlim || S->rsqstop <= 0 || S->rsq < S->rsqstop) S->done = 1;
    double a0 = S->rsq / oldrsq;
    for (int i = 0; i < S->nshift; i++) {
      double bi = b0, zni = 1, pai = a0;
      if (i > 0) {
        ms_coef(S, i, b0, a00, bo0, bi, zni);
        double c2 = S->z[i] * b0;
        pai = (c2 != 0.0) ? a0 * zni * bi / c2 : 0;
      }
      S->xb[i] = bi * S->s[i];
      S->pa[i] = pai;
      S->pzn[i] = zni;
      S->a[i] = pai;
      S->bo[i] = bi;
      S->zo[i] = S->z[i];
      S->z[i] = zni;
      if (S->drop && i > 0 && zni * zni * S->rsq < S->rsqstop) S->s[i] = 0;
    }
    S->xrun = 1;
  }
};

// r -= b_0 s_0 Ap; |r|^2 (lib/cgms.c:97-102).  Clears the previous
// iteration's x-update gate first, so k_ms_xp only runs after a FinMS.
template <int P>
__global__ void __launch_bounds__(256) k_ms_r(Field r, Field Ap, SolverDev *S, RedScratch rs) {
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  if (blockIdx.x == 0 && threadIdx.x == 0) S->xrun = 0;
  if (S->done) return;
  double pkp = (S->m2 * S->nrm[3] + S->nrm[0]) * S->s[0] * S->s[0];
  if (pkp <= 0) {   // lib/cgms.c:76: loss of precision, leave
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S->its += 1;
      S->done = 1;
    }
    return;
  }
  const double b0 = S->rsq / pkp;
  double red[1] = {0.0};
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < r.g->vh) {
    C rv[3], av[3];
    VT<P>::ld(r, i, rv);
    VT<P>::ld(Ap, i, av);
    R t0 = (R)(b0 * S->s[0]);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      rv[j].x = fma(-t0, av[j].x, rv[j].x);
      rv[j].y = fma(-t0, av[j].y, rv[j].y);
    }
    VT<P>::st(r, i, rv);
    C rr[3];
    VT<P>::ld(r, i, rr);   // the stored value (16-bit storage rounds)
#pragma unroll
    for (int j = 0; j < 3; j++) red[0] += (double)rr[j].x * rr[j].x + (double)rr[j].y * rr[j].y;
  }
  grid_partial<1>(red, rs);
}

// x_i += b_i s_i p_i (lib/cgms.c:91-96), then, unless the solve stopped this
// iteration, p_0 = a_0 p_0 + r and p_i = a_i p_i + zn_i r (lib/cgms.c:141-155):
// one pass over each shift's x and p.  Shifts with s_i = 0 are skipped.
template <int P>
__global__ void __launch_bounds__(256) k_ms_xp(MSPtrs V, Field r, SolverDev *S) {
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::S Sx;
  if (!S->xrun) return;
  const bool pupd = !S->done;
  const int64_t vp = r.vp;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.g->vh) return;
  C rv[3];
  if (pupd) VT<P>::ld(r, i, rv);
  for (int k = 0; k < S->nshift; k++) {
    if (k > 0 && S->s[k] == 0 && S->xb[k] == 0) continue;
    C xv[3], pv[3];
    VecLd<P>::ld((const Sx *)V.x[k], V.xn[k], vp, i, xv, (R)1);
    VecLd<P>::ld((const Sx *)V.p[k], V.pn[k], vp, i, pv, (R)1);
    const R t = (R)S->xb[k];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      xv[j].x = fma(t, pv[j].x, xv[j].x);
      xv[j].y = fma(t, pv[j].y, xv[j].y);
    }
    VecSt<P>::st((Sx *)V.x[k], V.xn[k], vp, i, xv);
    if (pupd && (k == 0 || S->s[k] != 0)) {
      R a = (R)S->pa[k], z = (R)(k == 0 ? 1.0 : S->pzn[k]);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        pv[j].x = fma(a, pv[j].x, z * rv[j].x);
        pv[j].y = fma(a, pv[j].y, z * rv[j].y);
      }
      VecSt<P>::st((Sx *)V.p[k], V.pn[k], vp, i, pv);
    }
  }
}

// ------------------------------------------------------------ host helpers
// rhs rs of FP64 colour-vector field src into rhs rd of dst (same lattice and
// parity; rhs-outer storage, so one contiguous copy)
static int rhs_copy(qamd_field *dst, int rd, const qamd_field *src, int rs) {
  const Field &d = dst->f, &s = src->f;
  if (d.prec != QAMD_F64 || s.prec != QAMD_F64 || d.ncomp != 3 || s.ncomp != 3 || d.vp != s.vp ||
      d.parity != s.parity || rd >= d.nrhs || rs >= s.nrhs) {
    set_error("rhs copy: field mismatch");
    return -1;
  }
  const size_t bytes = (size_t)3 * d.vp * sizeof(double2);
  QAMD_HIP(hipMemcpyAsync((char *)d.data + rd * bytes, (const char *)s.data + rs * bytes, bytes,
                          hipMemcpyDeviceToDevice, stream()));
  return 0;
}

struct State {
  SolverDev *S;    // device
  SolverDev *Sh;   // pinned host mirror
};

static void state_free(State &s) {
  dev_free(s.S);
  if (s.Sh) (void)hipHostFree(s.Sh);
  s.S = s.Sh = nullptr;
}

static int state_init(State &s, qamd_lattice *lat) {
  s.S = (SolverDev *)dev_alloc(sizeof(SolverDev));
  if (hipHostMalloc((void **)&s.Sh, sizeof(SolverDev), hipHostMallocDefault) != hipSuccess) s.Sh = nullptr;
  if (!s.S || !s.Sh) {
    if (!s.Sh) set_error("hipHostMalloc failed");
    state_free(s);
    return -1;
  }
  memset(s.Sh, 0, sizeof(SolverDev));
  return red_reserve((int)((lat->g.vh + 255) / 256));
}

static int push_state(State &s) {
  QAMD_HIP(hipMemcpyAsync(s.S, s.Sh, sizeof(SolverDev), hipMemcpyHostToDevice, stream()));
  return 0;
}
static int pull_state(State &s) {
  QAMD_HIP(hipMemcpyAsync(s.Sh, s.S, sizeof(SolverDev), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  return 0;
}

static inline dim3 grid1(const qamd_lattice *lat) { return dim3((unsigned)((lat->g.vh + 255) / 256)); }

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int check_vec(const qamd_field *f, int prec, int parity, const char *what) {
  if (!f || f->f.ncomp != 3 || f->f.nrhs != 1 || f->f.parity != parity || (prec >= 0 && f->f.prec != prec)) {
    set_error("%s: expected a single-rhs colour vector of parity %d%s", what, parity,
              prec >= 0 ? " at the required precision" : "");
    return -1;
  }
  return 0;
}

// Ap = m^2 p - D_eo D_oe p with the reductions of dslashp2
// (lib/stagDslash.c:546-575): |t|^2 -> nrm[0], |p|^2 -> nrm[3].
struct FinNormal {
  SolverDev *S;
  __device__ void operator()(const double *t) const {
    for (int i = 0; i < 4; i++) S->nrm[i] = t[i];
  }
};

static int apply_normal(const qamd_links *l, SolverDev *S, qamd_field *Ap, qamd_field *p, qamd_field *t, double m2,
                        const int *skip) {
  // partial sums of |t|^2 in set 1 and of |Ap|^2, |p|^2 in set 2, one finisher
  if (stag_dslash_launch(l, t->f, p->f, 0.0, nullptr, 1.0, nullptr, stream(), skip, 1, false)) return -1;
  if (stag_dslash_launch(l, Ap->f, t->f, m2, &p->f, -1.0, nullptr, stream(), skip, 2, false)) return -1;
  int nb = (int)((l->lat->g.vh + 255) / 256);
  hipLaunchKernelGGL((k_finish2<2, 2, FinNormal>), dim3(1), dim3(FT), 0, stream(), red_set(1), red_set(2), nb,
                     FinNormal{S}, skip);
  QAMD_HIP(hipGetLastError());
  return 0;
}

// p update of the plain CG (QLL's recurrences in mode 0)
static int cg_p(const qamd_lattice *lat, qamd_field *p, qamd_field *r, SolverDev *S) {
#define CGP(P) hipLaunchKernelGGL((k_cg_p<P>), grid1(lat), dim3(256), 0, stream(), p->f, r->f, S)
  switch (p->f.prec) {
    case QAMD_F64: CGP(QAMD_F64); break;
    case QAMD_F32: CGP(QAMD_F32); break;
    case QAMD_F16: CGP(QAMD_F16); break;
    default: CGP(QAMD_Q16); break;
  }
#undef CGP
  QAMD_HIP(hipGetLastError());
  return 0;
}

static int cg_xr(const qamd_lattice *lat, qamd_field *x, qamd_field *p, qamd_field *r, qamd_field *Ap, SolverDev *S) {
  int P = p->f.prec, PX = x->f.prec;
#define XR(A, B) hipLaunchKernelGGL((k_cg_xr<A, B>), grid1(lat), dim3(256), 0, stream(), x->f, p->f, r->f, Ap->f, S, red_set(0))
  if (P == QAMD_F64 && PX == QAMD_F64) XR(QAMD_F64, QAMD_F64);
  else if (P == QAMD_F32 && PX == QAMD_F32) XR(QAMD_F32, QAMD_F32);
  else if (P == QAMD_F16 && PX == QAMD_F32) XR(QAMD_F16, QAMD_F32);
  else if (P == QAMD_F16 && PX == QAMD_F16) XR(QAMD_F16, QAMD_F16);
  else {
    set_error("cg_xr: unsupported precision pair %d/%d", P, PX);
    return -1;
  }
#undef XR
  launch_finish<1>(red_set(0), (int)grid1(lat).x, FinCG{S}, stream(), &S->done);
  QAMD_HIP(hipGetLastError());
  return 0;
}

}  // namespace qamd

namespace qamd {
// 16-bit sloppy links are kept while the sloppy operator's absolute error
// stays below 0.9 m^2; measured on asqtad links, the inner CG needs
// <= 11 % more iterations than FP64 CG below 0.7 m^2 (+1 % at 0.83 m^2, the
// lightest shift of the 12-shift multishift test), 1.3-1.7x at 1.0-1.3 m^2
// and 4-5x at 1.6 m^2, against ~1.6x cheaper iterations.  When too coarse,
// *l32 receives FP32 links made from l64.  QAMD_SLOPPY_MAX overrides the
// factor (read once per process).
static double sloppy_max() {
  static double v = -1;
  if (v < 0) {
    const char *e = getenv("QAMD_SLOPPY_MAX");
    v = (e && atof(e) > 0) ? atof(e) : 0.9;
  }
  return v;
}
int sloppy_too_coarse(const qamd_links *l64, const qamd_links *ls, double mass, qamd_links **l32) {
  *l32 = nullptr;
  if (ls->prec != QAMD_F16 && ls->prec != QAMD_Q16) return 0;
  // the error of these links against these FP64 links at this mass is
  // measured once (a few Dslash launches) and kept with the sloppy links
  double eps;
  if (l64->ver && ls->ver && ls->err_ref_ver == l64->ver && ls->err_ver == ls->ver && ls->err_mass == mass) {
    eps = ls->err_eps;
  } else {
    if (qamd_stag_sloppy_error(l64, ls, mass, &eps)) return -1;
    ls->err_ref_ver = l64->ver;
    ls->err_ver = ls->ver;
    ls->err_mass = mass;
    ls->err_eps = eps;
  }
  if (eps < sloppy_max() * mass * mass) return 0;
  *l32 = qamd_links_create(l64->lat, l64->kind, QAMD_F32, l64->nterms);
  if (!*l32 || qamd_links_convert(*l32, l64)) {
    qamd_links_destroy(*l32);
    *l32 = nullptr;
    return -1;
  }
  return 0;
}
}  // namespace qamd

using namespace qamd;

extern "C" {

// QLL CG: all vectors at the links' precision, recurrences of cgSolveT
int qamd_stag_cg(const qamd_links *l, qamd_field *x, const qamd_field *b, double mass, double rsq, int maxits,
                 qamd_solve_info *info) {
  if (!l || !is_stag(l)) { set_error("stag_cg: staggered links required"); return -1; }
  if (l->prec != QAMD_F64 && l->prec != QAMD_F32) {
    set_error("stag_cg: links must be FP64 or FP32 (16-bit links run inside qamd_stag_cg_mixed)");
    return -1;
  }
  if (check_vec(x, l->prec, 0, "stag_cg x") || check_vec(b, l->prec, 0, "stag_cg b")) return -1;
  double t0 = now_s();
  qamd_lattice *lat = l->lat;
  State st;
  memset(&st, 0, sizeof st);
  qamd_field *t = qamd_field_create(lat, l->prec, 3, 1, 1);
  qamd_field *Ap = qamd_field_create(lat, l->prec, 3, 1, 0);
  qamd_field *r = qamd_field_create(lat, l->prec, 3, 1, 0);
  qamd_field *p = qamd_field_create(lat, l->prec, 3, 1, 0);
  int rc = -1;
  do {
    if (!t || !Ap || !r || !p || state_init(st, lat)) break;
    // x = 0, p = 0, r = b, rsq = |b|^2   (lib/cg.c:64-71)
    if (qamd_field_zero(x) || qamd_field_zero(p) || qamd_field_copy(r, b)) break;
    double bsq;
    if (qamd_field_norm2(r, &bsq)) break;
    SolverDev *Sh = st.Sh;
    Sh->rsq = bsq;
    Sh->m2 = mass * mass;
    Sh->rsqstop = rsq * bsq;
    Sh->itnlim = maxits > 0 ? maxits : 1;
    Sh->mode = 0;
    Sh->auto_stop = 1;
    if (bsq == 0) Sh->done = 1;   // zero source: the reference divides by zero; x = 0
    if (push_state(st)) break;
    int err = 0;
    while (!Sh->done) {
      for (int k = 0; k < BATCH && !err; k++) {
        err |= cg_p(lat, p, r, st.S);
        err |= apply_normal(l, st.S, Ap, p, t, mass * mass, &st.S->done);
        err |= cg_xr(lat, x, p, r, Ap, st.S);
      }
      if (err || pull_state(st)) {
        err = 1;
        break;
      }
    }
    if (err) break;
    if (info) {
      memset(info, 0, sizeof *info);
      info->iterations = Sh->its;
      info->iter_rsq = bsq > 0 ? Sh->rsq / bsq : 0;
      info->true_rsq = -1;
      info->converged = bsq == 0 || Sh->rsq < Sh->rsqstop;
      info->seconds = now_s() - t0;
    }
    rc = 0;
  } while (0);
  qamd_field_destroy(t);
  qamd_field_destroy(Ap);
  qamd_field_destroy(r);
  qamd_field_destroy(p);
  state_free(st);
  return rc;
}

// Mixed-precision CG with reliable updates.  The Krylov vectors r, p,
// s = Ap and the solution increment live in FP32; the Dslash reads the
// sloppy links (FP32, FP16 or Q16) and accumulates in FP32; the odd
// temporary t = D_oe r is kept in the links' format (16-bit with a per-site
// norm for 16-bit links); x, b and the true residual are FP64.  The inner
// iteration is the single-reduction (Chronopoulos-Gear) form of CG: one
// application of the operator to r and one global reduction per iteration,
// in two launches: t = D_oe r with |t|^2 partial sums, then the fused step,
// whose blocks form alpha and beta from the partial sums of |r|^2 and |t|^2,
// Ar = m^2 r - D_eo t (not stored), p = r + beta p, s = Ar + beta s,
// r -= alpha s with |r|^2 partial sums, and x += alpha p (two iterations at
// a time).  The recurrence for s = Ap assumes a linear operator: the links'
// rounding is fixed, and the one rounding applied to a vector (t) is at the
// 16-bit links' own level; restarts at every reliable update reset s.
// Reliable updates happen at batch boundaries when |r| has dropped below
// delta * max|r| since the last one.  The caller passes x = 0.
static int cg_mixed_core(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b,
                         double m2, double rsq, int maxits, double delta, qamd_solve_info *info) {
  double t0 = now_s();
  qamd_lattice *lat = l64->lat;
  const int ps = ls->prec;
  const qamd_links *lsc = ls;        // current sloppy links
  qamd_links *l32own = nullptr;      // FP32 links made after an FP16 stall
  int nesc = 0, nrestart = 0;
  State st;
  memset(&st, 0, sizeof st);
  qamd_field *r64 = qamd_field_create(lat, QAMD_F64, 3, 1, 0);
  qamd_field *Ax = qamd_field_create(lat, QAMD_F64, 3, 1, 0);
  qamd_field *t64 = qamd_field_create(lat, QAMD_F64, 3, 1, 1);
  qamd_field *r = qamd_field_create(lat, QAMD_F32, 3, 1, 0);
  qamd_field *p = qamd_field_create(lat, QAMD_F32, 3, 1, 0);
  qamd_field *sv = qamd_field_create(lat, QAMD_F32, 3, 1, 0);
  qamd_field *xs = qamd_field_create(lat, QAMD_F32, 3, 1, 0);
  // D_oe r is stored in the links' format (16-bit t with a per-site norm:
  // 8 bytes per site fewer written and gathered than FP32)
  qamd_field *t = qamd_field_create(lat, ps, 3, 1, 1);
  int rc = -1, nupd = 0, its = 0;
  const bool verbose = getenv("QAMD_VERBOSE") != nullptr;
  double bsq = 0, r2 = 0;
  do {
    if (!r64 || !Ax || !t64 || !r || !p || !sv || !xs || !t || state_init(st, lat)) break;
    if (qamd_field_norm2(b, &bsq)) break;
    if (bsq == 0) {
      if (qamd_field_zero(x)) break;
      rc = 0;
      break;
    }
    const double stop = rsq * bsq;
    SolverDev *Sh = st.Sh;
    Sh->m2 = m2;
    Sh->itnlim = 1 << 30;
    Sh->mode = 1;
    Sh->rsq_old = 1.0;
    if (push_state(st)) break;
    // true residual r64 = b - A x in FP64, r = r64 in FP32, |r64|^2 in
    // partial-sum set 0 and on the host
    // Iteration k (counted from the start, kit) reads the partial sums of
    // |r|^2 from set R[k & 1] and writes the next ones to R[1 - (k & 1)];
    // the D_oe launch leaves those of |t|^2 in set 1.
    const int nb1 = (int)grid1(lat).x;
    if (red_reserve(nb1)) break;
    const RedScratch rset[2] = {red_set(0), red_set(2)};
    long kit = 0;
    bool x_zero = true;   // the caller passes x = 0: the first residual is b
    auto reliable = [&](void) -> int {
      if (x_zero) {
        if (qamd_field_zero(Ax)) return -1;
        x_zero = false;
      } else if (apply_normal(l64, st.S, Ax, x, t64, m2, nullptr)) {
        return -1;
      }
      const RedScratch rr = rset[kit & 1];
      hipLaunchKernelGGL(k_ru_resid<QAMD_F32>, grid1(lat), dim3(256), 0, stream(), r64->f, b->f, Ax->f, r->f, st.S,
                         rr);
      launch_finish<1>(rr, nb1, FinTrue{st.S}, stream());
      QAMD_HIP(hipGetLastError());
      return pull_state(st);
    };
    // the next iteration takes beta = 0: p = r, s = A r
    auto restart = [&](void) -> int {
      Sh->cg.restart[kit & 1] = 1;
      Sh->cg.maxrr = sqrt(Sh->rsq);
      return push_state(st);
    };
    // the next iteration keeps p: r.s and p.s measured on the device
    auto keep = [&](void) -> int {
      hipLaunchKernelGGL(k_cg_keep_dots, grid1(lat), dim3(256), 0, stream(), r->f, p->f, sv->f, red_set(1));
      launch_finish<3>(red_set(1), nb1, FinKeep{st.S, (int)(kit & 1)}, stream());
      QAMD_HIP(hipGetLastError());
      return 0;
    };
    // QAMD_CG_KEEP=0: restart at every reliable update instead (measurements)
    const bool keep_p = !(getenv("QAMD_CG_KEEP") && atoi(getenv("QAMD_CG_KEEP")) == 0);
    if (reliable()) break;
    r2 = Sh->rsq;
    if (r2 < stop) {
      rc = 0;
      break;
    }
    if (qamd_field_zero(xs) || qamd_field_zero(p) || qamd_field_zero(sv) || restart()) break;
    int err = 0;
    // two launches per iteration: t = D_oe r (|t|^2 partial sums), then the
    // fused step, whose blocks form alpha and beta themselves; the step
    // updates x in pairs of iterations, so a batch must hold an even number
    static_assert(BATCH % 2 == 0, "the fused CG step updates x every second iteration");
    auto batch = [&]() -> int {
      int e = 0;
      for (int k = 0; k < BATCH && !e; k++, kit++) {
        e |= stag_dslash_launch(lsc, t->f, r->f, 0.0, nullptr, 1.0, nullptr, stream(), nullptr, 1, false);
        e |= stag_cgcg_step_launch(lsc, t->f, r->f, p->f, sv->f, xs->f, m2, &st.S->cg, (int)(kit & 1),
                                   rset[kit & 1], red_set(1), rset[1 - (kit & 1)], stream(), nullptr);
        e |= hipGetLastError() != hipSuccess;
      }
      return e;
    };
    while (its < maxits) {
      err |= batch();
      if (err || pull_state(st)) {
        err = 1;
        break;
      }
      its += BATCH;
      const CGScal &cg = Sh->cg;
      double rn = sqrt(cg.rsq);
      if (verbose)
        fprintf(stderr, "cg_mixed prec %d its %d rel|r| %.3e max %.3e rAr %.3e alpha %.3e beta %.3e\n", ps, its,
                rn / sqrt(bsq), cg.maxrr / sqrt(bsq), Sh->m2 * cg.rsq + cg.tt, cg.alpha, cg.beta);
      // update on a delta drop, on convergence, at the end, and when the
      // iterated residual grew well above the last true one (divergence)
      if (!isfinite(cg.rsq) || rn < delta * cg.maxrr || cg.rsq < stop || its >= maxits || cg.rsq > 4.0 * r2) {
        // reliable update: x += xs (xs = 0), r = b - A x in FP64
        hipLaunchKernelGGL(k_ru_accum<QAMD_F32>, grid1(lat), dim3(256), 0, stream(), x->f, xs->f);
        if (hipGetLastError() != hipSuccess || reliable()) {
          err = 1;
          break;
        }
        nupd++;
        const double r2prev = r2;
        r2 = Sh->rsq;
        if (verbose) fprintf(stderr, "cg_mixed reliable update %d: true rel|r| %.3e\n", nupd, sqrt(r2 / bsq));
        if (!isfinite(r2)) {
          set_error("cg_mixed: the solution is not finite after %d iterations", its);
          err = 1;
          break;
        }
        if (r2 < stop) break;
        // The sloppy operator's error shows up as drift between the iterated
        // and the true residual.  A large drift leaves p and s inconsistent
        // with the new r: restart the search direction.  When a whole cycle
        // did not reduce the true residual, the sloppy links are too coarse
        // for this operator: continue with FP32 links.
        const bool stalled = r2 > 0.25 * r2prev;
        if (stalled && lsc->prec != QAMD_F32) {
          if (!l32own) {
            l32own = qamd_links_create(lat, l64->kind, QAMD_F32, l64->nterms);
            if (!l32own || qamd_links_convert(l32own, l64)) {
              err = 1;
              break;
            }
          }
          lsc = l32own;
          if (t->f.prec != QAMD_F32) {
            qamd_field_destroy(t);
            if (!(t = qamd_field_create(lat, QAMD_F32, 3, 1, 1))) {
              err = 1;
              break;
            }
          }
          nesc++;
          if (verbose) fprintf(stderr, "cg_mixed: FP16 links stalled; continuing with FP32 links\n");
        }
        // Keep p across the reliable update.  The single-reduction
        // recurrences infer p.Ap from the previous step's scalars, which
        // assumes r orthogonal to the old p; the true residual is not (it
        // differs from the iterated one by the sloppy operator's drift), so
        // the next step measures r.s, p.s and r.p (k_cg_keep_dots) and takes
        // a conjugate beta and an exact line search, and the step after it
        // one more corrected step (cg_block_scalars).  Against restarting
        // (p = r) at every update, 32^4, QLL mass 0.02: Q16 links 1936 ->
        // 1872 iterations, FP32 links 1904 -> 1824.  Before new sloppy links
        // (escalation) the search restarts.
        if (keep_p && !stalled) {
          if (keep()) {
            err = 1;
            break;
          }
        } else {
          if (restart()) {
            err = 1;
            break;
          }
          nrestart++;
        }
      }
    }
    if (err) break;
    rc = 0;
  } while (0);
  if (info && rc == 0) {
    memset(info, 0, sizeof *info);
    info->iterations = its;
    info->reliable_updates = nupd;
    info->true_rsq = bsq > 0 ? r2 / bsq : 0;
    info->iter_rsq = info->true_rsq;
    info->converged = bsq == 0 || r2 < rsq * bsq;
    info->seconds = now_s() - t0;
    info->restarts = nrestart;
    info->escalations = nesc;
  }
  qamd_links_destroy(l32own);
  qamd_field_destroy(r64);
  qamd_field_destroy(Ax);
  qamd_field_destroy(t64);
  qamd_field_destroy(r);
  qamd_field_destroy(p);
  qamd_field_destroy(xs);
  qamd_field_destroy(sv);
  qamd_field_destroy(t);
  state_free(st);
  return rc;
}

int qamd_stag_sloppy_error(const qamd_links *l64, const qamd_links *ls, double mass, double *eps) {
  if (!l64 || !ls || !eps || !stag_pair(l64, ls) || l64->prec != QAMD_F64 || ls->prec == QAMD_F64 ||
      l64->lat->g.V != ls->lat->g.V) {
    set_error("stag_sloppy_error: need FP64 and sloppy staggered links on one lattice");
    return -1;
  }
  qamd_lattice *lat = l64->lat;
  const double m2 = mass * mass;
  qamd_field *v = qamd_field_create(lat, QAMD_F64, 3, 1, 0), *A = qamd_field_create(lat, QAMD_F64, 3, 1, 0);
  qamd_field *t = qamd_field_create(lat, QAMD_F64, 3, 1, 1), *vs = qamd_field_create(lat, ls->prec, 3, 1, 0);
  qamd_field *v32 = qamd_field_create(lat, QAMD_F32, 3, 1, 0), *ts = qamd_field_create(lat, ls->prec, 3, 1, 1);
  qamd_field *As = qamd_field_create(lat, QAMD_F32, 3, 1, 0), *d = qamd_field_create(lat, QAMD_F64, 3, 1, 0);
  int rc = -1;
  do {
    if (!v || !A || !t || !vs || !v32 || !ts || !As || !d) break;
    if (qamd_field_gaussian(v, 0x51077e, 3) || qamd_stag_normal(l64, A, v, t, mass, nullptr)) break;
    if (qamd_field_copy(vs, v) || qamd_field_copy(v32, v)) break;
    // the mixed solvers' sloppy operator: sloppy input copy and odd
    // temporary, FP32 output
    if (stag_dslash_launch(ls, ts->f, vs->f, 0.0, nullptr, 1.0, nullptr, stream())) break;
    if (stag_dslash_launch(ls, As->f, ts->f, m2, &v32->f, -1.0, nullptr, stream())) break;
    double d2, v2;
    if (qamd_field_copy(d, As) || qamd_field_axpby(-1.0, A, 1.0, d) || qamd_field_norm2(d, &d2) ||
        qamd_field_norm2(v, &v2))
      break;
    *eps = v2 > 0 ? sqrt(d2 / v2) : 0.0;
    rc = 0;
  } while (0);
  qamd_field_destroy(v);
  qamd_field_destroy(A);
  qamd_field_destroy(t);
  qamd_field_destroy(vs);
  qamd_field_destroy(v32);
  qamd_field_destroy(ts);
  qamd_field_destroy(As);
  qamd_field_destroy(d);
  return rc;
}

int qamd_stag_cg_mixed(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b,
                       double mass, double rsq, int maxits, double delta, qamd_solve_info *info) {
  if (!l64 || !ls || !stag_pair(l64, ls) || l64->prec != QAMD_F64 || ls->prec == QAMD_F64 ||
      l64->lat->g.V != ls->lat->g.V) {
    set_error("stag_cg_mixed: need FP64 links and FP32/FP16 sloppy links on one lattice");
    return -1;
  }
  if (check_vec(x, QAMD_F64, 0, "cg_mixed x") || check_vec(b, QAMD_F64, 0, "cg_mixed b")) return -1;
  if (delta <= 0 || delta >= 1) delta = 0.1;
  if (qamd_field_zero(x)) return -1;
  // 16-bit links whose operator error is not small against m^2 slow the
  // inner solve several-fold: start with FP32 links instead
  qamd_links *l32 = nullptr;
  if (sloppy_too_coarse(l64, ls, mass, &l32)) return -1;
  int rc = cg_mixed_core(l64, l32 ? l32 : ls, x, b, mass * mass, rsq, maxits, delta, info);
  if (l32 && info) info->escalations = 1;
  qamd_links_destroy(l32);
  return rc;
}

static int cgms_core(const qamd_links *l, qamd_field *const x[], const qamd_field *b, double mass,
                     const double shifts[], int n, double rsq, int maxits, qamd_solve_info *info, bool drop);

int qamd_stag_cgms(const qamd_links *l, qamd_field *const x[], const qamd_field *b, double mass,
                   const double shifts[], int n, double rsq, int maxits, qamd_solve_info *info) {
  return cgms_core(l, x, b, mass, shifts, n, rsq, maxits, info, false);
}

// QLL's multishift CG; with drop, converged shifts i > 0 stop updating
static int cgms_core(const qamd_links *l, qamd_field *const x[], const qamd_field *b, double mass,
                     const double shifts[], int n, double rsq, int maxits, qamd_solve_info *info, bool drop) {
  if (!l || !is_stag(l)) { set_error("stag_cgms: staggered links required"); return -1; }
  if (l->prec == QAMD_Q16) { set_error("stag_cgms: Q16 links are for the mixed-precision solvers only"); return -1; }
  if (n < 1 || n > MAXSHIFT) { set_error("stag_cgms: 1..%d shifts", MAXSHIFT); return -1; }
  if (check_vec(b, l->prec, 0, "stag_cgms b")) return -1;
  for (int i = 0; i < n; i++)
    if (check_vec(x[i], l->prec, 0, "stag_cgms x")) return -1;
  MallShareOff vectors_first;   // the 2n shift vectors fill the MALL
  double t0 = now_s();
  qamd_lattice *lat = l->lat;
  State st;
  memset(&st, 0, sizeof st);
  qamd_field *t = qamd_field_create(lat, l->prec, 3, 1, 1);
  qamd_field *Ap = qamd_field_create(lat, l->prec, 3, 1, 0);
  qamd_field *r = qamd_field_create(lat, l->prec, 3, 1, 0);
  qamd_field *ps[MAXSHIFT] = {nullptr};
  int rc = -1;
  do {
    int ok = t && Ap && r && !state_init(st, lat);
    for (int i = 0; i < n; i++) ok &= (ps[i] = qamd_field_create(lat, l->prec, 3, 1, 0)) != nullptr;
    if (!ok) break;
    double insq;
    if (qamd_field_norm2(b, &insq)) break;
    int err = 0;
    for (int i = 0; i < n; i++) err |= qamd_field_zero(x[i]);
    if (err) break;
    if (insq == 0) {   // lib/cgms.c:25-30
      if (info) memset(info, 0, sizeof *info);
      rc = 0;
      break;
    }
    if (qamd_field_copy(r, b)) break;
    for (int i = 0; i < n; i++) err |= qamd_field_copy(ps[i], r);
    if (err) break;
    SolverDev *Sh = st.Sh;
    Sh->nshift = n;
    for (int i = 0; i < n; i++) {
      Sh->shifts[i] = shifts[i];
      Sh->zo[i] = Sh->z[i] = 1;
      Sh->bo[i] = -1;
      Sh->a[i] = 0;
      Sh->s[i] = 1;
    }
    Sh->rsq = insq;
    Sh->rsqstop = rsq * insq;
    Sh->m2 = mass * mass;
    Sh->itnlim = maxits > 0 ? maxits : 1;
    Sh->xrun = 0;
    Sh->drop = drop ? 1 : 0;
    if (push_state(st)) break;
    MSPtrs V;
    memset(&V, 0, sizeof V);
    for (int i = 0; i < n; i++) {
      V.x[i] = x[i]->f.data;
      V.xn[i] = x[i]->f.norm;
      V.p[i] = ps[i]->f.data;
      V.pn[i] = ps[i]->f.norm;
    }
    while (!Sh->done && !err) {
      for (int k = 0; k < BATCH && !err; k++) {
        err |= apply_normal(l, st.S, Ap, ps[0], t, mass * mass, &st.S->done);
        RedScratch rs0 = red_set(0);
        int nb = (int)grid1(lat).x;
        switch (l->prec) {
          case QAMD_F64:
            hipLaunchKernelGGL(k_ms_r<QAMD_F64>, grid1(lat), dim3(256), 0, stream(), r->f, Ap->f, st.S, rs0);
            launch_finish<1>(rs0, nb, FinMS{st.S}, stream(), &st.S->done);
            hipLaunchKernelGGL(k_ms_xp<QAMD_F64>, grid1(lat), dim3(256), 0, stream(), V, r->f, st.S);
            break;
          case QAMD_F32:
            hipLaunchKernelGGL(k_ms_r<QAMD_F32>, grid1(lat), dim3(256), 0, stream(), r->f, Ap->f, st.S, rs0);
            launch_finish<1>(rs0, nb, FinMS{st.S}, stream(), &st.S->done);
            hipLaunchKernelGGL(k_ms_xp<QAMD_F32>, grid1(lat), dim3(256), 0, stream(), V, r->f, st.S);
            break;
          default:
            hipLaunchKernelGGL(k_ms_r<QAMD_F16>, grid1(lat), dim3(256), 0, stream(), r->f, Ap->f, st.S, rs0);
            launch_finish<1>(rs0, nb, FinMS{st.S}, stream(), &st.S->done);
            hipLaunchKernelGGL(k_ms_xp<QAMD_F16>, grid1(lat), dim3(256), 0, stream(), V, r->f, st.S);
            break;
        }
        if (hipGetLastError() != hipSuccess) err = 1;
      }
      if (err || pull_state(st)) err = 1;
    }
    if (err) break;
    if (info) {
      memset(info, 0, sizeof *info);
      info->iterations = Sh->its;
      info->iter_rsq = Sh->rsq / insq;
      info->true_rsq = -1;
      info->converged = Sh->rsq < Sh->rsqstop;
      info->seconds = now_s() - t0;
    }
    rc = 0;
  } while (0);
  for (int i = 0; i < n; i++) qamd_field_destroy(ps[i]);
  qamd_field_destroy(t);
  qamd_field_destroy(Ap);
  qamd_field_destroy(r);
  state_free(st);
  return rc;
}

int qamd_stag_cgms_mixed(const qamd_links *l64, const qamd_links *ls, qamd_field *const x[], const qamd_field *b,
                         double mass, const double shifts[], int n, double rsq, int maxits, qamd_solve_info *info) {
  if (!l64 || !ls || l64->prec != QAMD_F64 || ls->prec == QAMD_F64 || l64->kind != 0 || ls->kind != 0 ||
      l64->nterms != ls->nterms || l64->lat->g.V != ls->lat->g.V) {
    set_error("stag_cgms_mixed: need FP64 and sloppy doublestored staggered links (the refinement runs the "
              "block kernels)");
    return -1;
  }
  if (n < 1 || n > MAXSHIFT) { set_error("stag_cgms_mixed: 1..%d shifts", MAXSHIFT); return -1; }
  for (int i = 0; i < n; i++)
    if (check_vec(x[i], QAMD_F64, 0, "cgms_mixed x")) return -1;
  if (check_vec(b, QAMD_F64, 0, "cgms_mixed b")) return -1;
  qamd_lattice *lat = l64->lat;
  // the multishift recurrences run in FP32 (FP16 links are promoted for them)
  const qamd_links *l32 = ls;
  qamd_links *own = nullptr;
  if (ls->prec != QAMD_F32) {
    own = qamd_links_create(lat, l64->kind, QAMD_F32, l64->nterms);
    if (!own || qamd_links_convert(own, l64)) {
      qamd_links_destroy(own);
      return -1;
    }
    l32 = own;
  }
  qamd_field *bs = qamd_field_create(lat, QAMD_F32, 3, 1, 0);
  qamd_field *xs[MAXSHIFT] = {nullptr};
  int rc = -1;
  do {
    if (!bs || qamd_field_copy(bs, b)) break;
    int ok = 1;
    for (int i = 0; i < n; i++) ok &= (xs[i] = qamd_field_create(lat, QAMD_F32, 3, 1, 0)) != nullptr;
    if (!ok) break;
    qamd_solve_info msinfo;
    // FP32 recurrences do not resolve residuals far below 1e-6; refine
    double ms_rsq = rsq > 1e-12 ? rsq : 1e-12;
    // (a shift whose FP32 residual is there stops updating: its refinement
    // starts from it)
    if (cgms_core(l32, xs, bs, mass, shifts, n, ms_rsq, maxits, &msinfo, true)) break;
    int err = 0;
    // refinement: the shifts, lightest first, in blocks of up to 8 that
    // share every link load (each rhs with its own mass^2 + shift, starting
    // from its FP32 solution); a block runs on the 16-bit links when they are
    // accurate enough for every shift in it (see sloppy_too_coarse), else on
    // FP32 links
    double eps = 0;
    if ((ls->prec == QAMD_F16 || ls->prec == QAMD_Q16) && qamd_stag_sloppy_error(l64, ls, mass, &eps)) break;
    for (int i = 0; i < n && !err; i++) err |= qamd_field_copy(x[i], xs[i]);
    int idx[MAXSHIFT];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++)   // insertion sort by shift
      for (int j = i; j > 0 && shifts[idx[j]] < shifts[idx[j - 1]]; j--) {
        const int t = idx[j];
        idx[j] = idx[j - 1];
        idx[j - 1] = t;
      }
    for (int c0 = 0; c0 < n && !err; c0 += 8) {
      const int kb = n - c0 < 8 ? n - c0 : 8;
      bool coarse = false;
      for (int q = 0; q < kb; q++) coarse |= eps >= sloppy_max() * (mass * mass + shifts[idx[c0 + q]]);
      const qamd_links *lg = coarse ? l32 : ls;
      qamd_field *X = qamd_field_create(lat, QAMD_F64, 3, kb, 0);
      qamd_field *B = qamd_field_create(lat, QAMD_F64, 3, kb, 0);
      double m2r[8];
      qamd_solve_info ri[8];
      err = !X || !B;
      for (int q = 0; q < kb && !err; q++) {
        m2r[q] = mass * mass + shifts[idx[c0 + q]];
        err |= rhs_copy(X, q, x[idx[c0 + q]], 0) || rhs_copy(B, q, b, 0);
      }
      err = err || stag_cg_block_core(l64, lg, X, B, m2r, true, rsq, maxits, 0.1, ri);
      for (int q = 0; q < kb && !err; q++) {
        const int i = idx[c0 + q];
        err |= rhs_copy(x[i], 0, X, q);
        if (info) {
          info[i] = ri[q];
          info[i].iterations += msinfo.iterations;
          info[i].escalations = (lg != ls);
          info[i].seconds += (c0 == 0 && q == 0) ? msinfo.seconds : 0;
        }
      }
      qamd_field_destroy(X);
      qamd_field_destroy(B);
    }
    if (err) break;
    rc = 0;
  } while (0);
  for (int i = 0; i < n; i++) qamd_field_destroy(xs[i]);
  qamd_field_destroy(bs);
  qamd_links_destroy(own);
  return rc;
}

}  // extern "C"
