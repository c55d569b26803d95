/* solvers.h -- CG and multishift CG of the QLL C API (drop-in for
 * include/solvers.h).  Vectors are host arrays of nelem*VLEN reals; the
 * recurrences run in libqll_amd's HIP kernels. */
#ifndef QLL_SOLVERS_H
#define QLL_SOLVERS_H
#include "thread.h"

typedef struct {
  int adj;
  int wantAnorm2;
  double Anorm2;
} LinopInfo;
#define LINOP_INFO_DEFAULT {0, 0, 0}

#define real float
#define PR(x) x##F
#include "solversP.h"
#undef real
#undef PR

#define real double
#define PR(x) x##D
#include "solversP.h"
#undef real
#undef PR
#endif
