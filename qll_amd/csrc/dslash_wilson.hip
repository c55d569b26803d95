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
// One thread per (site, rhs); with several right-hand sides (rhs fastest in
// memory) the lanes of a wavefront that share a site load the same link
// elements, so each link is fetched once for all of them.
#include "qamd_internal.h"
#include "qamd_objects.h"
#include "qamd_kernels.h"

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
};

template <class C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <class C> __device__ __forceinline__ C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <class C> __device__ __forceinline__ C cmi(C a) { C r; r.x = a.y; r.y = -a.x; return r; }   // -i a
template <class C> __device__ __forceinline__ C cpi(C a) { C r; r.x = -a.y; r.y = a.x; return r; }   // +i a
template <class C> __device__ __forceinline__ C cneg(C a) { a.x = -a.x; a.y = -a.y; return a; }

template <int P> struct SpIO;
template <> struct SpIO<QAMD_F64> {
  __device__ static double2 ld(const Field &f, int64_t i, float) { return ((const double2 *)f.data)[i]; }
  __device__ static void st(const Field &f, int64_t i, double2 v) { ((double2 *)f.data)[i] = v; }
};
template <> struct SpIO<QAMD_F32> {
  __device__ static float2 ld(const Field &f, int64_t i, float) { return ((const float2 *)f.data)[i]; }
  __device__ static void st(const Field &f, int64_t i, float2 v) { ((float2 *)f.data)[i] = v; }
};
template <> struct SpIO<QAMD_F16> {
  __device__ static float2 ld(const Field &f, int64_t i, float n) {
    float2 v = __half22float2(((const __half2 *)f.data)[i]);
    v.x *= n;
    v.y *= n;
    return v;
  }
  __device__ static void st(const Field &f, int64_t i, float2 v) { ((__half2 *)f.data)[i] = __floats2half2_rn(v.x, v.y); }
};
template <> struct SpIO<QAMD_Q16> {
  __device__ static float2 ld(const Field &f, int64_t i, float n) {
    short2 s = ((const short2 *)f.data)[i];
    n *= 1.0f / Q16_ONE;
    return make_float2(s.x * n, s.y * n);
  }
  __device__ static void st(const Field &f, int64_t i, float2 v) { ((short2 *)f.data)[i] = q16_pack(v.x, v.y); }
};

struct FinStore1 {
  double *dst;
  __device__ void operator()(const double *t) const { dst[0] = t[0]; }
};

template <int P, bool RED>
__global__ void __launch_bounds__(256) k_wilson(WArgs<P> A) {
  typedef typename Prec<P>::R R;
  typedef typename Prec<P>::C C;
  typedef typename Prec<P>::S S;
  if (A.skip && *A.skip) return;
  // local copies: the kernel reads its argument structs only here
  const Field fin = A.in, fw = A.w;
  const Field fout = A.out;
  const Geom g = A.g;
  const int nr = fin.nrhs;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int cb = (int)(gid / nr), r = (int)(gid - (int64_t)cb * nr);
  double red[1] = {0.0};
  if (cb < g.vh) {
    Site s = site_of(g, cb, A.par);
    constexpr int G = LinkG<S>::G;
    const S *lp = link_base<S, 8>(A.links, cb);
    C acc[4][3];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[a][c].x = acc[a][c].y = (R)0;
    // one link group (G terms, one 16-byte load per element) per iteration;
    // the iterations are not unrolled, which keeps the spinor working set of
    // one group in registers (unrolled, the hoisted loads of all eight terms
    // spill)
#pragma unroll 1
    for (int tg = 0; tg < 8 / G; tg++) {
    S q[9][G];
#pragma unroll
    for (int e = 0; e < 9; e++) ntload_group(lp + link_off<S>(tg * G, e), q[e]);
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
      const int t = tg * G + gg;
      const int mu = t >> 1;
      const bool bwd = t & 1;
      int nb = nbr_cb_rt(g, s, mu, bwd ? -1 : 1);
      float nrm = 1.0f;
      if constexpr (P == QAMD_F16 || P == QAMD_Q16) nrm = fin.norm[fin.nidx(r, nb)] * A.lscale[t];
      C psi[4][3];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) psi[a][c] = SpIO<P>::ld(fin, fin.cidx(r, 3 * a + c, nb), nrm);
      // h = psi_up + s A psi_dn, s = +1 for backward terms
      C h[2][3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C d0, d1;   // A psi_dn
        if (mu == 0) { d0 = cpi(psi[3][c]); d1 = cpi(psi[2][c]); }
        else if (mu == 1) { d0 = cneg(psi[3][c]); d1 = psi[2][c]; }
        else if (mu == 2) { d0 = cpi(psi[2][c]); d1 = cmi(psi[3][c]); }
        else { d0 = psi[2][c]; d1 = psi[3][c]; }
        h[0][c] = bwd ? cadd(psi[0][c], d0) : csub(psi[0][c], d0);
        h[1][c] = bwd ? cadd(psi[1][c], d1) : csub(psi[1][c], d1);
      }
      C m[9];
#pragma unroll
      for (int e = 0; e < 9; e++) m[e] = to_c(q[e][gg], (R)0);
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
        C l0, l1;   // A^+ chi
        if (mu == 0) { l0 = cmi(uh[1][c]); l1 = cmi(uh[0][c]); }
        else if (mu == 1) { l0 = uh[1][c]; l1 = cneg(uh[0][c]); }
        else if (mu == 2) { l0 = cmi(uh[0][c]); l1 = cpi(uh[1][c]); }
        else { l0 = uh[0][c]; l1 = uh[1][c]; }
        acc[2][c] = bwd ? cadd(acc[2][c], l0) : csub(acc[2][c], l0);
        acc[3][c] = bwd ? cadd(acc[3][c], l1) : csub(acc[3][c], l1);
      }
    }
    }
    const R b = (R)A.b, a = (R)A.a;
    C o[4][3];
    float wn = 1.0f;
    if constexpr (P == QAMD_F16 || P == QAMD_Q16)
      if (A.has_w) wn = fw.norm[fw.nidx(r, cb)];
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C v;
        v.x = b * acc[sp][c].x;
        v.y = b * acc[sp][c].y;
        if (A.has_w) {
          C wv = SpIO<P>::ld(fw, fw.cidx(r, 3 * sp + c, cb), wn);
          v.x = fma(a, wv.x, v.x);
          v.y = fma(a, wv.y, v.y);
        }
        o[sp][c] = v;
      }
    float on = 1.0f;
    constexpr bool NORMED = (P == QAMD_F16 || P == QAMD_Q16);
    if (NORMED) {
      float mx = 0.f;
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int c = 0; c < 3; c++) mx = fmaxf(mx, fmaxf(fabsf((float)o[sp][c].x), fabsf((float)o[sp][c].y)));
      fout.norm[fout.nidx(r, cb)] = mx;
      on = mx > 0.f ? 1.0f / mx : 0.f;
    }
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C v = o[sp][c];
        if (RED) red[0] += (double)v.x * v.x + (double)v.y * v.y;
        if (NORMED) {
          v.x *= on;
          v.y *= on;
        }
        SpIO<P>::st(fout, fout.cidx(r, 3 * sp + c, cb), v);
      }
  }
  if (RED) grid_partial<1>(red, A.rs);
}

template <int P>
static int wlaunch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                   double *red_out, hipStream_t st, const int *skip) {
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
  for (int t = 0; t < 8; t++) A.lscale[t] = (P == QAMD_Q16) ? l->scale[t] / Q16_ONE : l->scale[t];
  int64_t nthreads = A.g.vh * in.nrhs;
  int nb = (int)((nthreads + 255) / 256);
  if (red_out) {
    if (red_reserve(nb)) return -1;
    A.rs = red_set(0);
    A.red_out = red_out;
    hipLaunchKernelGGL((k_wilson<P, true>), dim3(nb), dim3(256), 0, st, A);
    launch_finish<1>(A.rs, nb, FinStore1{red_out}, st, skip);
  } else {
    hipLaunchKernelGGL((k_wilson<P, false>), dim3(nb), dim3(256), 0, st, A);
  }
  QAMD_HIP(hipGetLastError());
  return 0;
}

int wilson_dslash_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                         double *red_out, hipStream_t st, const int *skip) {
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
    case QAMD_F64: return wlaunch<QAMD_F64>(l, out, in, a, w, b, red_out, st, skip);
    case QAMD_F32: return wlaunch<QAMD_F32>(l, out, in, a, w, b, red_out, st, skip);
    case QAMD_F16: return wlaunch<QAMD_F16>(l, out, in, a, w, b, red_out, st, skip);
    default: return wlaunch<QAMD_Q16>(l, out, in, a, w, b, red_out, st, skip);
  }
}

}  // namespace qamd
