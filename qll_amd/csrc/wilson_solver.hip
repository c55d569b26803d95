// Wilson even-odd solver: CG on the normal equations of
//   Mhat = 1 - kappa^2 H_eo H_oe,   Mhat^+ = g5 Mhat g5,
// for a block of right-hand sides that share every link load, inside FP64
// defect correction when a sloppy (FP32/FP16) link set is given.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_site.h"
#include "qamd_kernels.h"

#include <string.h>
#include <math.h>
#include <chrono>

namespace qamd {

constexpr int MAXRHS = 16;

__global__ void __launch_bounds__(256) k_gamma5(Field f) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int cb = (int)(g / f.nrhs), r = (int)(g % f.nrhs);
  if (cb >= f.g->vh) return;
  double2 v[12];
  site_load<12>(f, r, cb, v);
#pragma unroll
  for (int j = 6; j < 12; j++) v[j] = make_double2(-v[j].x, -v[j].y);
  site_store<12>(f, r, cb, v);
}

struct FinRhs {
  double *dst;
  int n;
  __device__ void operator()(const double *t) const {
    for (int i = 0; i < n; i++) dst[i] = t[i];
  }
};

// per-rhs |f|^2 (mode 0) or Re <f, h> (mode 1)
__global__ void __launch_bounds__(256) k_rhs_dot(Field f, Field h, int mode, RedScratch rs) {
  double acc[MAXRHS];
#pragma unroll
  for (int i = 0; i < MAXRHS; i++) acc[i] = 0.0;
  int64_t n = f.g->vh * (int64_t)f.nrhs;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    int cb = (int)(g / f.nrhs), r = (int)(g % f.nrhs);
    double2 a[12], b[12];
    site_load<12>(f, r, cb, a);
    if (mode) site_load<12>(h, r, cb, b);
    double s = 0;
#pragma unroll
    for (int j = 0; j < 12; j++) s += mode ? a[j].x * b[j].x + a[j].y * b[j].y : a[j].x * a[j].x + a[j].y * a[j].y;
#pragma unroll
    for (int i = 0; i < MAXRHS; i++)
      if (i == r) acc[i] += s;
  }
  grid_partial<MAXRHS>(acc, rs);
}

// e_r += alpha_r p_r; p_r = r_r + beta_r p_r (one pass; alpha at a, beta at b)
__global__ void __launch_bounds__(256) k_rhs_ep(Field e, Field p, Field r, const double *a, const double *b) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int cb = (int)(g / p.nrhs), q = (int)(g % p.nrhs);
  if (cb >= p.g->vh) return;
  double2 ve[12], vp[12], vr[12];
  const double al = a[q], be = b[q];
  site_load<12>(e, q, cb, ve);
  site_load<12>(p, q, cb, vp);
  site_load<12>(r, q, cb, vr);
#pragma unroll
  for (int j = 0; j < 12; j++) {
    ve[j] = make_double2(ve[j].x + al * vp[j].x, ve[j].y + al * vp[j].y);
    vp[j] = make_double2(vr[j].x + be * vp[j].x, vr[j].y + be * vp[j].y);
  }
  site_store<12>(e, q, cb, ve);
  site_store<12>(p, q, cb, vp);
}

struct WS {   // work space of one precision
  qamd_lattice *lat;
  int prec, nrhs;
  qamd_field *r, *p, *q, *s, *t;   // t odd, others even
  double *coef;                        // device: a[16], b[16], norms[16]
  double *coefh;                       // pinned host mirror
};

static void ws_free(WS &w) {
  qamd_field_destroy(w.r);
  qamd_field_destroy(w.p);
  qamd_field_destroy(w.q);
  qamd_field_destroy(w.s);
  qamd_field_destroy(w.t);
  dev_free(w.coef);
  if (w.coefh) (void)hipHostFree(w.coefh);
  memset(&w, 0, sizeof w);
}

static int ws_init(WS &w, qamd_lattice *lat, int prec, int nrhs) {
  memset(&w, 0, sizeof w);
  w.lat = lat;
  w.prec = prec;
  w.nrhs = nrhs;
  w.r = qamd_field_create(lat, prec, 12, nrhs, 0);
  w.p = qamd_field_create(lat, prec, 12, nrhs, 0);
  w.q = qamd_field_create(lat, prec, 12, nrhs, 0);
  w.s = qamd_field_create(lat, prec, 12, nrhs, 0);
  w.t = qamd_field_create(lat, prec, 12, nrhs, 1);
  w.coef = (double *)dev_alloc(64 * sizeof(double));
  if (hipHostMalloc((void **)&w.coefh, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) w.coefh = nullptr;
  if (!w.r || !w.p || !w.q || !w.s || !w.t || !w.coef || !w.coefh) {
    ws_free(w);
    return -1;
  }
  return 0;
}

static inline unsigned nblk(const qamd_field *f) {
  return (unsigned)((f->lat->g.vh * f->f.nrhs + 255) / 256);
}

static int rhs_dot(WS &w, const qamd_field *f, const qamd_field *h, double *out) {
  int nb = 1024;
  if (red_reserve(nb)) return -1;
  RedScratch rs = red_set(0);
  hipLaunchKernelGGL(k_rhs_dot, dim3(nb), dim3(256), 0, stream(), f->f, h ? h->f : f->f, h ? 1 : 0, rs);
  launch_finish<MAXRHS>(rs, nb, FinRhs{w.coef + 32, f->f.nrhs}, stream());
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipMemcpyAsync(w.coefh + 32, w.coef + 32, f->f.nrhs * sizeof(double), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  memcpy(out, w.coefh + 32, f->f.nrhs * sizeof(double));
  return 0;
}

// out = Mhat in, or Mhat^+ in = g5 Mhat g5 in = in - kappa^2 (g5 H_eo g5)(g5 H_oe g5) in,
// whose hopping terms are the Dslash with every projector sign flipped
static int apply_mhat(const qamd_links *l, WS &w, qamd_field *out, const qamd_field *in, double kappa, int dagger) {
  const bool dag = dagger != 0;
  if (wilson_dslash_launch(l, w.t->f, in->f, 0.0, nullptr, 1.0, nullptr, stream(), nullptr, dag)) return -1;
  if (wilson_dslash_launch(l, out->f, w.t->f, 1.0, &in->f, -kappa * kappa, nullptr, stream(), nullptr, dag))
    return -1;
  return 0;
}

// per-rhs sums of slot r over nb blocks of partial-sum set 0 -> host
static int rhs_sums_to_host(WS &w, int nb, double *out) {
  launch_finish<MAXRHS>(red_set(0), nb, FinRhs{w.coef + 32, w.nrhs}, stream());
  QAMD_HIP(hipGetLastError());
  QAMD_HIP(hipMemcpyAsync(w.coefh + 32, w.coef + 32, w.nrhs * sizeof(double), hipMemcpyDeviceToHost, stream()));
  QAMD_HIP(hipStreamSynchronize(stream()));
  memcpy(out, w.coefh + 32, w.nrhs * sizeof(double));
  return 0;
}

// coefficients a[16], b[16] to the device (the stream is idle when called)
static int coef_up(WS &w, const double *a, const double *b) {
  memcpy(w.coefh, a, w.nrhs * sizeof(double));
  memcpy(w.coefh + 16, b, w.nrhs * sizeof(double));
  QAMD_HIP(hipMemcpyAsync(w.coef, w.coefh, 32 * sizeof(double), hipMemcpyHostToDevice, stream()));
  return 0;
}

// CGNE from x = 0 on Mhat e = rhs, to |r_normal|^2 <= tol2 |r0_normal|^2 per
// rhs.  Per iteration: q = Mhat p with |q_r|^2 in the second Dslash's
// epilogue; alpha_r = |r_r|^2 / |q_r|^2; r_r -= alpha_r Mhat^+ q_r in the
// epilogue of Mhat^+'s second Dslash (Mhat^+ q not stored) with |r_r|^2;
// beta_r; one pass e_r += alpha_r p_r, p_r = r_r + beta_r p_r.
static int cgne_inner(const qamd_links *l, WS &w, qamd_field *e, const qamd_field *rhs, double kappa, double tol2,
                      int maxits, int *its_out) {
  const int n = rhs->f.nrhs;
  const int nb = (int)nblk(rhs);
  const double k2 = -kappa * kappa;
  double rr[MAXRHS], rr0[MAXRHS], qq[MAXRHS], rrn[MAXRHS], al[MAXRHS], be[MAXRHS];
  int done[MAXRHS];
  if (qamd_field_zero(e)) return -1;
  if (apply_mhat(l, w, w.r, rhs, kappa, 1)) return -1;
  if (qamd_field_copy(w.p, w.r)) return -1;
  if (rhs_dot(w, w.r, nullptr, rr)) return -1;
  for (int i = 0; i < n; i++) {
    rr0[i] = rr[i];
    done[i] = rr[i] == 0;
  }
  int it = 0;
  while (it < maxits) {
    int all = 1;
    for (int i = 0; i < n; i++) all &= done[i];
    if (all) break;
    // q = Mhat p, |q_r|^2
    if (wilson_dslash_launch(l, w.t->f, w.p->f, 0.0, nullptr, 1.0, nullptr, stream(), nullptr, false)) return -1;
    if (wilson_dslash_rhs_launch(l, w.q->f, w.t->f, 1.0, &w.p->f, k2, stream(), false, 1, nullptr, nullptr,
                                 red_set(0)))
      return -1;
    if (rhs_sums_to_host(w, nb, qq)) return -1;
    for (int i = 0; i < n; i++) {
      al[i] = (!done[i] && qq[i] > 0) ? rr[i] / qq[i] : 0.0;
      be[i] = 0.0;
    }
    if (coef_up(w, al, be)) return -1;
    // r -= alpha Mhat^+ q, |r_r|^2
    if (wilson_dslash_launch(l, w.t->f, w.q->f, 0.0, nullptr, 1.0, nullptr, stream(), nullptr, true)) return -1;
    if (wilson_dslash_rhs_launch(l, w.s->f, w.t->f, 1.0, &w.q->f, k2, stream(), true, 2, &w.r->f, w.coef,
                                 red_set(0)))
      return -1;
    if (rhs_sums_to_host(w, nb, rrn)) return -1;
    it++;
    for (int i = 0; i < n; i++) {
      if (!done[i] && rrn[i] <= tol2 * rr0[i]) done[i] = 1;
      be[i] = (!done[i] && rr[i] > 0) ? rrn[i] / rr[i] : 0.0;
      rr[i] = rrn[i];
    }
    if (coef_up(w, al, be)) return -1;
    hipLaunchKernelGGL(k_rhs_ep, dim3(nb), dim3(256), 0, stream(), e->f, w.p->f, w.r->f, (const double *)w.coef,
                       (const double *)(w.coef + 16));
    QAMD_HIP(hipGetLastError());
  }
  *its_out = it;
  return 0;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace qamd

using namespace qamd;

extern "C" {

int qamd_field_gamma5(qamd_field *f) {
  if (!f || f->f.ncomp != 12) {
    set_error("gamma5: spinor field required");
    return -1;
  }
  hipLaunchKernelGGL(k_gamma5, dim3(nblk(f)), dim3(256), 0, stream(), f->f);
  QAMD_HIP(hipGetLastError());
  return 0;
}

int qamd_wilson_cgne(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b, double kappa,
                     double rsq, int maxits, qamd_solve_info *info) {
  if (!l64 || l64->kind != 1 || (ls && (ls->kind != 1 || ls->lat->g.V != l64->lat->g.V))) {
    set_error("wilson_cgne: Wilson links required");
    return -1;
  }
  const int n = b->f.nrhs;
  if (n < 1 || n > MAXRHS || x->f.nrhs != n || x->f.ncomp != 12 || b->f.ncomp != 12 || x->f.parity != 0 ||
      b->f.parity != 0 || x->f.prec != l64->prec || b->f.prec != l64->prec) {
    set_error("wilson_cgne: x and b must be even-parity spinor blocks (<= %d rhs) at the links' precision", MAXRHS);
    return -1;
  }
  double t0 = now_s();
  qamd_lattice *lat = l64->lat;
  const qamd_links *li = ls ? ls : l64;
  WS w64, wi;
  memset(&wi, 0, sizeof wi);
  if (ws_init(w64, lat, l64->prec, n)) return -1;
  if (li != l64 && ws_init(wi, lat, li->prec, n)) {
    ws_free(w64);
    return -1;
  }
  WS &win = (li != l64) ? wi : w64;
  qamd_field *res = qamd_field_create(lat, l64->prec, 12, n, 0);
  qamd_field *rs = (li != l64) ? qamd_field_create(lat, li->prec, 12, n, 0) : nullptr;
  qamd_field *es = (li != l64) ? qamd_field_create(lat, li->prec, 12, n, 0) : nullptr;
  qamd_field *e64 = qamd_field_create(lat, l64->prec, 12, n, 0);
  int rc = -1, its = 0, cycles = 0;
  double bb[MAXRHS], rn[MAXRHS];
  do {
    if (!res || !e64 || (li != l64 && (!rs || !es))) break;
    if (qamd_field_zero(x) || rhs_dot(w64, b, nullptr, bb)) break;
    // inner tolerance per cycle: limited by the inner precision
    double tol2 = li->prec == QAMD_F64 ? 1e-12 : li->prec == QAMD_F32 ? 1e-10 : 1e-5;
    int err = 0;
    while (!err) {
      // res = b - Mhat x in the outer precision
      if (apply_mhat(l64, w64, res, x, kappa, 0)) { err = 1; break; }
      if (qamd_field_axpby(1.0, b, -1.0, res) || rhs_dot(w64, res, nullptr, rn)) { err = 1; break; }
      int conv = 1;
      for (int i = 0; i < n; i++) conv &= (bb[i] == 0 || rn[i] <= rsq * bb[i]);
      if (conv || its >= maxits) break;
      int it = 0;
      if (li != l64) {
        if (qamd_field_copy(rs, res) || cgne_inner(li, win, es, rs, kappa, tol2, maxits - its, &it) ||
            qamd_field_copy(e64, es)) { err = 1; break; }
      } else {
        if (cgne_inner(l64, w64, e64, res, kappa, tol2, maxits - its, &it)) { err = 1; break; }
      }
      its += it;
      cycles++;
      if (qamd_field_axpby(1.0, e64, 1.0, x)) { err = 1; break; }
      if (it == 0) break;
    }
    if (err) break;
    if (info)
      for (int i = 0; i < n; i++) {
        memset(&info[i], 0, sizeof info[i]);
        info[i].iterations = its;
        info[i].reliable_updates = cycles;
        info[i].true_rsq = bb[i] > 0 ? rn[i] / bb[i] : 0;
        info[i].iter_rsq = info[i].true_rsq;
        info[i].converged = bb[i] == 0 || rn[i] <= rsq * bb[i];
        info[i].seconds = now_s() - t0;
      }
    rc = 0;
  } while (0);
  qamd_field_destroy(res);
  qamd_field_destroy(rs);
  qamd_field_destroy(es);
  qamd_field_destroy(e64);
  ws_free(w64);
  if (li != l64) ws_free(wi);
  return rc;
}

}  // extern "C"
