/* Internal declarations shared by the C files of the QLL API. */
#ifndef QLL_HOST_H
#define QLL_HOST_H
#include "qll.h"
#include "qll_amd.h"

extern int qllVlen[2];
void qllPartitionGeom(int *lx, int *nx, const int *x, int n, int nd, int dist);
void qllToDevLayout(const Layout *l, qamd_qll_layout *q);
void qllFatal(const char *what);
void qllRegPut(const void *key, void *obj);
void *qllRegGet(const void *key);
void *qllRegTake(const void *key);
qamd_lattice *qllLattice(const Layout *l);
#endif
