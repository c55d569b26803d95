/*
 * thread.h -- OpenMP team context of the QLL C API (drop-in for
 * include/thread.h of QLL, lines 20-80: ThreadCtx, BEGIN_PARALLEL,
 * END_PARALLEL, TI0/TI1, setOff*).  Kept so that user Linop callbacks written
 * against QLL compile and run unchanged; the device work of libqll_amd never
 * runs inside these teams.
 */
#ifndef QLL_THREAD_H
#define QLL_THREAD_H
#ifdef _OPENMP
#include <omp.h>
#define OMP_PARALLEL _Pragma("omp parallel")
#define OMP_NUMTHREADS omp_get_num_threads()
#define OMP_THREADNUM omp_get_thread_num()
#define OMP_BARRIER _Pragma("omp barrier")
#else
#define OMP_PARALLEL
#define OMP_NUMTHREADS 1
#define OMP_THREADNUM 0
#define OMP_BARRIER
#endif
#define NUMTHREADS OMP_NUMTHREADS
#define THREADNUM OMP_THREADNUM
#define TBARRIER OMP_BARRIER
#ifndef TIBDEFAULT
#define TIBDEFAULT 0
#endif
#define MAXTHREADS 512

/* same member order and types as QLL's ThreadCtx (thread.h:31-38) */
typedef struct {
  double **sum;
  int volatile *tbar;
  int tid;
  int nid;
  int tib;
  int n0, n1, d0;
} ThreadCtx;

#define BEGIN_PARALLEL                 \
  {                                    \
    double *_tsum[MAXTHREADS];         \
    int volatile _tbar[MAXTHREADS];    \
    OMP_PARALLEL                       \
    {                                  \
      ThreadCtx TC0, *TC = &TC0;       \
      TC0.tid = OMP_THREADNUM;         \
      TC0.nid = OMP_NUMTHREADS;        \
      TC0.tbar = _tbar;                \
      TC0.sum = _tsum;                 \
      TC0.tib = TIBDEFAULT;            \
      TC0.n0 = TC0.n1 = TC0.d0 = 0;    \
      _tbar[TC0.tid] = 0;              \
      OMP_BARRIER;
#define END_PARALLEL \
  }                  \
  }

#define SUBSET_DEFS int ti0, ti1, off, len
#define SUBSET_DEFS2 int ti0, ti1
#define TI0(n) ((TC->tid * (n)) / TC->nid)
#define TI1(n) (((TC->tid + 1) * (n)) / TC->nid)
#define setOff2(n) ti0 = TI0(n); ti1 = TI1(n)
#define setOffSub2(s) setOff2((s).endOuter - (s).beginOuter); ti0 += (s).beginOuter; ti1 += (s).beginOuter

/* global sum over the team (lib/common.c:52-72, single rank) */
void gsum(double *x, int n, ThreadCtx *TC);
#endif
