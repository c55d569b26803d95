/*
 * thread.h -- OpenMP team context of the QLL C API (drop-in for
 * include/thread.h of QLL: ThreadCtx, BEGIN_PARALLEL/END_PARALLEL, the
 * thread-range macros TI0/TI1 and their "C" variants, setOff*, SET_TIB,
 * T0WAIT and TWAIT0, include/thread.h:20-80 and its twait0).  User Linop
 * callbacks written against QLL (e.g. in the style of QLL's dslashp2, which
 * partitions with setOffSubC2 and sums with gsum) compile and run unchanged;
 * the device work of libqll_amd never runs inside these teams.
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

/* same member order and types as QLL's ThreadCtx (thread.h:31-38).
 * tib, n0, n1, d0 describe the "C" partition: the range is cut into
 * 10*nid parts, the first tib of them go to no thread (so thread 0 gets
 * less work while it also drives communication in QLL), and thread t takes
 * parts [n0, n1) of the remaining d0. */
typedef struct {
  double **sum;
  int volatile *tbar;
  int tid;
  int nid;
  int tib;
  int n0, n1, d0;
} ThreadCtx;

/* set the "C" partition of the current team member TC */
#define SET_TIB(n)                           \
  do {                                       \
    TC->tib = (n);                           \
    TC->n0 = 10 * TC->tid - TC->tib;         \
    if (TC->n0 < 0) TC->n0 = 0;              \
    TC->n1 = 10 * (TC->tid + 1) - TC->tib;   \
    TC->d0 = 10 * TC->nid - TC->tib;         \
  } while (0)

#define DEFINE_THREAD_CTX          \
  ThreadCtx TC0, *TC = &TC0;       \
  TC0.tid = OMP_THREADNUM;         \
  TC0.nid = OMP_NUMTHREADS;        \
  TC0.tbar = _tbar;                \
  TC0.sum = _tsum;                 \
  SET_TIB(TIBDEFAULT)

#define BEGIN_PARALLEL                 \
  {                                    \
    double *_tsum[MAXTHREADS];         \
    int volatile _tbar[MAXTHREADS];    \
    OMP_PARALLEL                       \
    {                                  \
      DEFINE_THREAD_CTX;               \
      _tbar[TC0.tid] = 0;              \
      OMP_BARRIER;
#define END_PARALLEL \
  }                  \
  }

#define SUBSET_DEFS int ti0, ti1, off, len
#define SUBSET_DEFS2 int ti0, ti1
#define SET_SUBSET(s)
/* even split of n items over the team */
#define TI0(n) ((TC->tid * (n)) / TC->nid)
#define TI1(n) (((TC->tid + 1) * (n)) / TC->nid)
/* split with the first tib tenths-per-thread held back (SET_TIB) */
#define TI0C(n) ((TC->n0 * (n)) / TC->d0)
#define TI1C(n) ((TC->n1 * (n)) / TC->d0)
#define setOff(n) ti0 = TI0(n); ti1 = TI1(n); off = ti0; len = ti1 - ti0
#define setOffC(n) ti0 = TI0C(n); ti1 = TI1C(n); off = ti0; len = ti1 - ti0
#define setOffSub(s)                                                     \
  setOff((s).endOuter - (s).beginOuter); ti0 += (s).beginOuter;          \
  ti1 += (s).beginOuter; off += (s).beginOuter
#define setOffSubC(s)                                                    \
  setOffC((s).endOuter - (s).beginOuter); ti0 += (s).beginOuter;         \
  ti1 += (s).beginOuter; off += (s).beginOuter
#define setOff2(n) ti0 = TI0(n); ti1 = TI1(n)
#define setOffC2(n) ti0 = TI0C(n); ti1 = TI1C(n)
#define setOffSub2(s) setOff2((s).endOuter - (s).beginOuter); ti0 += (s).beginOuter; ti1 += (s).beginOuter
#define setOffSubC2(s) setOffC2((s).endOuter - (s).beginOuter); ti0 += (s).beginOuter; ti1 += (s).beginOuter

#define SYNC __sync_synchronize()

/* every thread waits for all others (QLL's T0WAIT is the team barrier) */
#define T0WAIT OMP_BARRIER

/* the team waits for thread 0 to arrive; thread 0 does not wait.  Each
 * thread counts its arrivals in tbar[tid]; thread t > 0 proceeds once thread
 * 0 has arrived as often as t has. */
static inline void twait0(ThreadCtx *TC)
{
  int volatile *cnt = TC->tbar;
  SYNC;
  int mine = ++cnt[TC->tid];
  if (TC->tid != 0)
    while (cnt[0] < mine) {
    }
  SYNC;
}
#define TWAIT0 twait0(TC)

/* global sum over the team (lib/common.c:52-72, single rank) */
void gsum(double *x, int n, ThreadCtx *TC);
#endif
