/*
 * qll.h -- the QLL C API, drop-in for QLL's include/qll.h, served by
 * libqll_amd on one MI355X (gfx950).
 *
 * Every QLL function exists twice, suffix F (float) and D (double), exactly
 * as in QLL (include/qll.h:17-27).  Fields cross the API in QLL's host layout
 * (the Layout built by stagDslashSetup); the library converts them to its
 * device layout and runs all arithmetic in HIP kernels.  As in QLL,
 * createStaggeredLinks/createWilsonLinks take the caller's link arrays; they
 * are copied to the GPU at that call, so call create* again after changing
 * the links.
 */
#ifndef QLL_H
#define QLL_H
#include <stddef.h>
#include "layout.h"
#include "thread.h"
#include "solvers.h"

#ifdef __cplusplus
extern "C" {
#endif

/* lib/common.c:9-28: 64-byte aligned host allocation */
void *myalloc(size_t size);
/* the rank used by printf0; set by stagDslashSetup (lib/common.c:6) */
extern int myrank;

typedef struct {
  double one_link;
  double three_staple;
  double five_staple;
  double seven_staple;
  double lepage;
  double naik;
} Fat7Coeffs;
#define FAT7COEFFS_ZERO (Fat7Coeffs){0, 0, 0, 0, 0, 0}

#define real float
#define PR(x) x##F
#include "qllP.h"
#undef real
#undef PR

#define real double
#define PR(x) x##D
#include "qllP.h"
#undef real
#undef PR

/* VLEN (sites per SIMD vector) used by stagDslashSetup to build the host
 * layout; QLL fixes it at compile time (include/vec_types.h:1-52).  Default
 * 1; any power of two the lattice supports is accepted. */
void qllSetVlen(int vlenF, int vlenD);

#ifdef __cplusplus
}
#endif

/* Generic names (include/qll.h:29-58): Field, fieldNew, transporterDo, ...
 * resolve to the F or D function through P(x), which vla.h (included by
 * common.h) selects with -DDOUBLE; the team macros open an OpenMP team with
 * a subset's outer-site range split over the threads (ti0, ti1). */
#ifndef QLL_NO_GENERIC_NAMES
#define beginParallel BEGIN_PARALLEL
#define beginParallelSubset(s) \
  beginParallel {              \
    SUBSET_DEFS2;              \
    setOffSub2(s);
#define endParallel END_PARALLEL
#define endParallelSubset \
  }                       \
  endParallel
#define singleThread _Pragma("omp single")

#define Field P(Field)
#define fieldNew P(fieldNew)
#define fieldFree P(fieldFree)
#define fieldSet P(fieldSet)
#define fieldZero P(fieldZero)
#define fieldSumElem P(fieldSumElem)
#define fieldNorm2 P(fieldNorm2)

#define Transporter P(Transporter)
#define transporterNew P(transporterNew)
#define transporterFree P(transporterFree)
#define transporterStartT P(transporterStartT)
#define transporterLocalT P(transporterLocalT)
#define transporterWaitT P(transporterWaitT)
#define transporterDoT P(transporterDoT)
#define transporterDo P(transporterDo)

#define gaugeDoublestore P(gaugeDoublestore)
#define gaugePlaq P(gaugePlaq)
#define smearFat7 P(smearFat7)

/* set every real of field ff on the thread's range to r; its squared norm */
#define F_eq_r(ff, r) X_eq_r((ff)->f, (r), (ff)->nelem, ti0, ti1)
#define lnorm2_F(ff) norm2_X((ff)->f, (ff)->nelem, ti0, ti1)
#define sumDoubleArray(x, n) gsum(x, n, TC)
#endif
#endif
