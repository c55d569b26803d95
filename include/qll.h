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
#endif
