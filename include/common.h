/*
 * common.h -- QLL's program-level header (drop-in for include/common.h of
 * QLL) without QMP: libqll_amd runs one lattice on one GPU, so the rank is
 * always 0 and no message-passing type or call is needed.  QLL programs such
 * as its test drivers include this header for:
 *   - the precision selector P(x), `real` (vla.h, chosen by -DDOUBLE), qll.h
 *     with its generic names, and the NAIK / SHIFT13 switches;
 *   - printf0 and the other rank-0 print/trace macros (common.h:12-27);
 *   - the growable-array helpers ARRAY_CREATE/GROW/APPEND/COPY/CLONE
 *     (common.h:29-49);
 *   - the shift-table builders makeShift, makeShiftSub, makeShiftMulti,
 *     makeShiftMultiSub (common.h:88-91; lib/shifts.c:497-736), served by
 *     libqll_amd;
 *   - the prototypes of the qop* entry points (common.h:66-86).  These are
 *     defined by the program's QOPQDP wrapper (qopstag.c, qopwilson.c), not
 *     by QLL or libqll_amd.
 * Not provided: QLL's message-plan internals (makeGather, startSend,
 * prepareShiftBuf, ..., common.h:51-53,92-110) -- on one GPU neighbours are
 * computed on the device and nothing is sent.  A program that also calls
 * QMP itself (QLL's drivers call QMP_get_number_of_nodes, QMP_barrier,
 * QMP_sum_double) needs QMP or a one-rank stand-in for those calls.
 */
#ifndef QLL_COMMON_H
#define QLL_COMMON_H
#include <stdio.h>
#include <stdlib.h>

#include "vla.h"
#include "qll.h"

#define NAIK
#define SHIFT13

#define if0 if (myrank == 0)
#define printf0(...)            \
  if (myrank == 0) {            \
    printf(__VA_ARGS__);        \
    fflush(stdout);             \
  }
#define TRACE                                                     \
  do {                                                            \
    if0 {                                                         \
      printf("%s %s %i\n", __FILE__, __func__, __LINE__);         \
      fflush(stdout);                                             \
    }                                                             \
  } while (0)
#define TRACET                                                              \
  do {                                                                      \
    if0 {                                                                   \
      printf("%i: %s %s %i\n", TC->tid, __FILE__, __func__, __LINE__);      \
      fflush(stdout);                                                       \
    }                                                                       \
  } while (0)
#define TRACE_ALL                                                        \
  do {                                                                   \
    printf("%i %s %s %i\n", myrank, __FILE__, __func__, __LINE__);      \
    fflush(stdout);                                                      \
  } while (0)
#define BEGIN_FATAL (void)0
#define PRINT_FATAL(...) printf0(__VA_ARGS__)
#define END_FATAL exit(-1)

/* print label s, then the n values of v with format f */
#define PRINTV(s, f, v, n)                               \
  do {                                                   \
    printf(s);                                           \
    for (int _k = 0; _k < (n); _k++) printf(" " f, (v)[_k]); \
    printf("\n");                                        \
  } while (0)
#define PRINTV0(s, f, v, n) \
  do {                      \
    if0 {                   \
      PRINTV(s, f, v, n);   \
    }                       \
  } while (0)

/* growable array nm of type tp: count n##nm, capacity nm##len */
#define ARRAY_CREATE(tp, nm) \
  int CAT(n, nm) = 0, CAT(nm, len) = 0; \
  tp *nm = NULL
#define ARRAY_GROW(tp, nm, ln)                              \
  if (CAT(n, nm) + (ln) > CAT(nm, len)) {                   \
    CAT(nm, len) += (ln);                                   \
    nm = (tp *)realloc(nm, CAT(nm, len) * sizeof(tp));      \
  }
#define ARRAY_APPEND(tp, nm, vl)                              \
  if (CAT(n, nm) == CAT(nm, len)) {                           \
    CAT(nm, len) = CAT(nm, len) ? 2 * CAT(nm, len) : 16;      \
    nm = (tp *)realloc(nm, CAT(nm, len) * sizeof(tp));        \
  }                                                           \
  nm[CAT(n, nm)++] = (vl)
#define ARRAY_COPY(tp, ds, nm)                                  \
  do {                                                          \
    for (int _k = 0; _k < CAT(n, nm); _k++) (ds)[_k] = (nm)[_k]; \
  } while (0)
#define ARRAY_CLONE(tp, ds, nm)                     \
  do {                                              \
    (ds) = (tp *)myalloc(CAT(n, nm) * sizeof(tp));  \
    ARRAY_COPY(tp, ds, nm);                         \
  } while (0)

#ifdef __cplusplus
extern "C" {
#endif

/* shift tables for arbitrary displacements (lib/shifts.c:497-736): si[n]
 * describes dst(x) = src(x - disp[n]) on the sites of subs[n] ("all",
 * "even", "odd"); pidx/sidx give each outer site's source outer site (-1
 * outside the subset; -(i)-2 in pidx when the SIMD lanes arrive permuted by
 * perm), no message members are set.  Use with transporterNew for a
 * displacement other than the getShift set; free with freeShift. */
void makeShift(ShiftIndices *si, Layout *l, int *disp);
void makeShiftSub(ShiftIndices *si, Layout *l, int *disp, char *sub);
void makeShiftMulti(ShiftIndices *si[], Layout *l, int *disp[], int ndisp);
void makeShiftMultiSub(ShiftIndices *si[], Layout *l, int *disp[], char *subs[], int ndisp);
/* releases what makeShift* allocated in si (no counterpart in QLL, which
 * keeps its tables for the program's lifetime) */
void freeShift(ShiftIndices *si);

/* the program's QOPQDP wrapper (qopstag.c:6-459, qopwilson.c:6-402) */
void qopStagDslashInit(int *argc, char ***argv);
void qopStagDslashFini(void);
void qopStagDslashSetup(Layout *l, int *rg);
void qopShift(Layout *l, real *xx, real *yy, int nelem, int dir, int len);
void qopTransportM(Layout *l, real *xx, real *mm, real *yy, int dir, int len);
void qopSmear(P(Field) *fl[], P(Field) *ll[], Fat7Coeffs *coef, P(Field) *u[], P(Field) *ul[]);
void qopDslash(Layout *l, real *x, real *u[8], real *u3[8], real mass, real *y, char *sub);
void qopSolve(Layout *l, real *x, real *u[8], real *u3[8], real mass, real *y, double rsq, char *sub);
void qopSolveMulti(Layout *l, real *x[], real *u[8], real *u3[8], double masses[], real *y, int nmasses,
                   double rsq, char *sub);
void qopWilsonDslashInit(int *argc, char ***argv);
void qopWilsonDslashFini(void);
void qopWilsonDslashSetup(Layout *l);
void qopSproj(Layout *l, real *x, real *y, int dir, int sign);
void qopSrecon(Layout *l, real *x, real *y, int dir, int sign);
void qopWilsonDslash(Layout *l, real *x, real *u[8], real mass, int sign, real *y, char *sub);
void qopWilsonSolve(Layout *l, real *x, real *u[8], real mass, real *y, double rsq, char *sub);
void qopWilsonSolveMulti(Layout *l, real *x[], real *u[8], double masses[], real *y, int nmasses, double rsq,
                         char *sub);

#ifdef __cplusplus
}
#endif
#endif
