/*
 * vla.h -- the part of QLL's vector-op header (include/vla.h) that QLL
 * programs use directly: the precision selector P(x) and the element type
 * `real` chosen by -DDOUBLE (include/vla.h:4-8, include/vec_types.h:257-265),
 * and the host BLAS-1 helpers behind the generic names of qll.h (F_eq_r,
 * lnorm2_F) and of QLL's test drivers (X_vpeq_r_times_X, makeU).
 *
 * These helpers work on the caller's host arrays in QLL layout, called by
 * every thread of an OpenMP team on its own range [i0, i1) of outer sites
 * (nelem reals per site and SIMD lane, so reals [nelem*VLEN*i0,
 * nelem*VLEN*i1) with VLEN the layout's nSitesInner); they are plain host
 * loops, as in QLL, and run no device work.  QLL's SIMD kernels
 * (V_vpeq_M_times_xVp, the spin projectors, pack/blend, ...) are internal to
 * its Dslash and are replaced by the HIP kernels of libqll_amd, so they are
 * not declared here.
 */
#ifndef QLL_VLA_H
#define QLL_VLA_H

#ifndef CAT
#define CAT(a, b) QLL_CAT_(a, b)
#define QLL_CAT_(a, b) a##b
#endif

/* P pastes without expanding its argument first, so P(name) also works when
 * name is itself one of the generic-name macros below or in qll.h */
#ifdef DOUBLE
typedef double real;
#define P(x) x##D
#else
typedef float real;
#define P(x) x##F
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define QLL_VLA_DECL(T, S)                                                              \
  double sumElem_X##S(T *vv, int nelem, int i0, int i1);                                \
  double norm2_X##S(T *vv, int nelem, int i0, int i1);                                  \
  void r_veq_norm2_X##S(double *r, T *vv, int nelem, int i0, int i1);                   \
  double re_X_dot_X##S(T *xx, T *yy, int nelem, int i0, int i1);                        \
  void r_veq_re_X_dot_X##S(double *r, T *vv, T *ww, int nelem, int i0, int i1);         \
  void X_veq_r##S(T *rr, T aa, int nelem, int i0, int i1);                              \
  void X_veq_r_times_X##S(T *rr, T aa, T *vv, int nelem, int i0, int i1);               \
  void X_vpeq_r_times_X##S(T *rr, T aa, T *vv, int nelem, int i0, int i1);              \
  void X_veq_X_plus_r_times_X##S(T *rr, T *vv, T aa, T *ww, int nelem, int i0, int i1); \
  void makeU##S(T *mm, int nc, int i0, int i1);
QLL_VLA_DECL(float, F)
QLL_VLA_DECL(double, D)
#undef QLL_VLA_DECL

#ifdef __cplusplus
}
#endif

/* generic names (include/vla.h:120-140) */
#define X_eq_r X_veq_r
#define X_veq_r P(X_veq_r)
#define X_veq_r_times_X P(X_veq_r_times_X)
#define X_vpeq_r_times_X P(X_vpeq_r_times_X)
#define r_veq_norm2_X P(r_veq_norm2_X)
#define sumElem_X P(sumElem_X)
#define norm2_X P(norm2_X)
#define re_X_dot_X P(re_X_dot_X)
#define r_veq_re_X_dot_X P(r_veq_re_X_dot_X)
#define X_veq_X_plus_r_times_X P(X_veq_X_plus_r_times_X)
#define V_eq_zero(x) X_veq_r(x, 0, 1, ti0, ti1)
#define V_eq_V(x, y) X_veq_r_times_X(x, 1, y, 1, ti0, ti1)
#define r_eq_norm2_V(r, x) r_veq_norm2_X(r, x, 1, ti0, ti1)
#define r_eq_re_V_dot_V(r, x, y) r_veq_re_X_dot_X(r, x, y, 1, ti0, ti1)
#define V_peq_r_times_V(x, a, y) X_veq_X_plus_r_times_X(x, x, a, y, 1, ti0, ti1)
#define V_meq_r_times_V(x, a, y) X_veq_X_plus_r_times_X(x, x, -(a), y, 1, ti0, ti1)
#endif
