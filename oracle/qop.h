/*
 * qop* entry points of QLL's test programs, served by the FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  In QLL these functions (declared at
 * include/common.h:66-86, defined in qopstag.c and qopwilson.c) wrap QOPQDP,
 * the reference's own source of ground truth; here they compute the same
 * quantities with the FP64 CPU oracle (oracle.c) on host arrays in QLL's
 * layout, so that a test program written like QLL's testdslash.c and
 * testwilson.c checks libqll_amd against them.  The product never links this.
 *
 * As in QLL, the names are unsuffixed and the library is built once per
 * precision: -DDOUBLE gives real = double (libqopD.so), otherwise float
 * (libqopF.so).  Arithmetic is FP64 in both.
 */
#ifndef QLL_QOP_ORACLE_H
#define QLL_QOP_ORACLE_H
#include "qll.h"

#ifdef DOUBLE
#define QOP_REAL double
#define QOP_FIELD FieldD
#else
#define QOP_REAL float
#define QOP_FIELD FieldF
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* staggered (qopstag.c:6-459) */
void qopStagDslashInit(int *argc, char ***argv);
void qopStagDslashFini(void);
void qopStagDslashSetup(Layout *l, int *rg);
void qopShift(Layout *l, QOP_REAL *xx, QOP_REAL *yy, int nelem, int dir, int len);
void qopTransportM(Layout *l, QOP_REAL *xx, QOP_REAL *mm, QOP_REAL *yy, int dir, int len);
void qopSmear(QOP_FIELD *fl[], QOP_FIELD *ll[], Fat7Coeffs *coef, QOP_FIELD *u[], QOP_FIELD *ul[]);
void qopDslash(Layout *l, QOP_REAL *x, QOP_REAL *u[8], QOP_REAL *u3[8], QOP_REAL mass, QOP_REAL *y,
               char *sub);
void qopSolve(Layout *l, QOP_REAL *x, QOP_REAL *u[8], QOP_REAL *u3[8], QOP_REAL mass, QOP_REAL *y,
              double rsq, char *sub);
void qopSolveMulti(Layout *l, QOP_REAL *x[], QOP_REAL *u[8], QOP_REAL *u3[8], double masses[],
                   QOP_REAL *y, int nmasses, double rsq, char *sub);

/* Wilson (qopwilson.c:6-402) */
void qopWilsonDslashInit(int *argc, char ***argv);
void qopWilsonDslashFini(void);
void qopWilsonDslashSetup(Layout *l);
void qopSproj(Layout *l, QOP_REAL *x, QOP_REAL *y, int dir, int sign);
void qopSrecon(Layout *l, QOP_REAL *x, QOP_REAL *y, int dir, int sign);
void qopWilsonDslash(Layout *l, QOP_REAL *x, QOP_REAL *u[8], QOP_REAL mass, int sign, QOP_REAL *y,
                     char *sub);
void qopWilsonSolve(Layout *l, QOP_REAL *x, QOP_REAL *u[8], QOP_REAL mass, QOP_REAL *y, double rsq,
                    char *sub);
void qopWilsonSolveMulti(Layout *l, QOP_REAL *x[], QOP_REAL *u[8], double masses[], QOP_REAL *y,
                         int nmasses, double rsq, char *sub);

#ifdef __cplusplus
}
#endif
#endif
