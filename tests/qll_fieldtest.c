/*
 * QLL's fieldtest.c checks (fieldtest.c:43-225) and two more API paths,
 * written the way a QLL program is: against common.h with the generic names
 * of qll.h (Field, fieldNew, beginParallelSubset, F_eq_r, lnorm2_F,
 * sumDoubleArray, singleThread, transporter*, gaugeDoublestore) and P(x)
 * chosen by -DDOUBLE, linked to libqll_amd and to the qop* entry points of
 * the FP64 oracle (test infrastructure in place of QOPQDP).
 *
 *  1. makeU gives unitary matrices; forward transport of v1 with U and
 *     getShift(dir, +1), then backward with the doublestored adjoint and
 *     getShift(dir, -1), returns v1 (fieldtest.c:183-191); the forward result
 *     equals U(x) v1(x+dir) built from qopShift.
 *  2. matrix transport (td = 2) equals qopTransportM (fieldtest.c:193-204);
 *     td = 0 is a plain shift, equal to qopShift.
 *  3. transporterStartT / LocalT / WaitT inside a team on the even subset
 *     write the even sites only.
 *  4. makeShift for displacement (1,-2,0,3): the transporter shifts by it
 *     (td = 0) and transports with U (td = 2, accumulating), against composed
 *     qopShift calls; the site table points at the right source sites.
 *  5. a Linop in the style of QLL's dslashp2 (thread 0 applies D_oe, D_eo
 *     with dslash2; the team forms m^2 x - D_eo D_oe x on its setOffSubC2
 *     range and sums with gsum) passed to cgSolve gives solve2's solution.
 *
 * usage: qll_fieldtest <vlen> Lx Ly Lz Lt    (exit status: failures)
 */
#include <math.h>
#include <string.h>

#include "common.h"

#ifdef DOUBLE
#define TOL 1e-12
#define TOL_SOLVE 1e-8
#define SOLVE_RSQ 1e-22
#else
#define TOL 1e-5
#define TOL_SOLVE 1e-3
#define SOLVE_RSQ 1e-10
#endif

static Layout layout;
static int nfail;

static void check(const char *what, double err, double tol)
{
  int ok = err <= tol && !isnan(err);
  printf("%s %-44s %.3e (tol %.1e)\n", ok ? "ok  " : "FAIL", what, err, tol);
  nfail += !ok;
}

/* relative difference over n reals */
static double reldiff(const real *a, const real *b, long n)
{
  double d = 0, w = 0;
  for (long i = 0; i < n; i++) {
    d += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
    w += (double)b[i] * b[i];
  }
  return sqrt(d / (w > 0 ? w : 1));
}

/* deterministic values in [-1, 1), different per (tag, index) */
static real val(long i, int tag)
{
  unsigned long h = (unsigned long)i * 2654435761ul + (unsigned long)tag * 40503ul + 12345ul;
  h ^= h >> 13;
  h *= 0x5bd1e995ul;
  h ^= h >> 15;
  return (real)((double)(h % 2000003ul) / 1000001.5 - 1.0);
}

/* dest(x) = m(x) * v(x) for colour vectors in QLL layout (column-major m) */
static void mat_vec(real *dest, const real *m, const real *v)
{
  int NI = layout.nSitesInner;
  for (long o = 0; o < layout.nSitesOuter; o++)
    for (int ii = 0; ii < NI; ii++) {
      const real *mo = m + 18L * NI * o + ii, *vo = v + 6L * NI * o + ii;
      real *d = dest + 6L * NI * o + ii;
      for (int r = 0; r < 3; r++) {
        double re = 0, im = 0;
        for (int c = 0; c < 3; c++) {
          double mr = mo[NI * 2 * (r + 3 * c)], mi = mo[NI * (2 * (r + 3 * c) + 1)];
          double vr = vo[NI * 2 * c], vi = vo[NI * (2 * c + 1)];
          re += mr * vr - mi * vi;
          im += mr * vi + mi * vr;
        }
        d[NI * 2 * r] = (real)re;
        d[NI * (2 * r + 1)] = (real)im;
      }
    }
}

typedef struct {
  P(StaggeredLinks) *sle, *slo;
  real *tmp;
  double m2;
  Subset sube, subo;
} LinArgs;

/* A x = m^2 x - D_eo D_oe x on the even sites, Anorm2 = m^2|x|^2 + |D_oe x|^2
 * (the operator of QLL's dslashp2, lib/stagDslash.c:546-575) */
static void normalLinop(real *Ax, real *x, LinopInfo *li, void *args, ThreadCtx *TC)
{
  LinArgs *a = (LinArgs *)args;
  TBARRIER;
  if (TC->tid == 0) {
    P(dslash2)(a->slo, a->tmp, 0, x);
    P(dslash2)(a->sle, Ax, 0, a->tmp);
  }
  TBARRIER;
  SUBSET_DEFS2;
  int NI = layout.nSitesInner;
  double r[2] = {0, 0};
  setOffSubC2(a->sube);
  for (long i = 6L * NI * ti0; i < 6L * NI * ti1; i++) {
    Ax[i] = (real)(a->m2 * x[i] - Ax[i]);
    r[0] += (double)x[i] * x[i];
  }
  setOffSubC2(a->subo);
  for (long i = 6L * NI * ti0; i < 6L * NI * ti1; i++) r[1] += (double)a->tmp[i] * a->tmp[i];
  gsum(r, 2, TC);
  li->Anorm2 = a->m2 * r[0] + r[1];
}

int main(int argc, char *argv[])
{
  if (argc < 6) {
    printf("usage: %s vlen Lx Ly Lz Lt\n", argv[0]);
    return 2;
  }
  int vlen = atoi(argv[1]);
  int ls[4];
  for (int i = 0; i < 4; i++) ls[i] = atoi(argv[2 + i]);
#ifdef DOUBLE
  qllSetVlen(1, vlen);
#else
  qllSetVlen(vlen, 1);
#endif
  layout.nranks = 1;
  layout.myrank = 0;
  P(stagDslashSetup)(&layout, 4, ls, NULL);
  qopStagDslashSetup(&layout, NULL);
  int vol = layout.nSites, nc = 3, nelemV = 2 * nc, nelemM = 2 * nc * nc;
  int NI = layout.nSitesInner;
  printf0("fieldtest: %s, vlen %d, %d %d %d %d, %d threads\n",
#ifdef DOUBLE
          "D",
#else
          "F",
#endif
          NI, ls[0], ls[1], ls[2], ls[3], omp_get_max_threads());

  Field v1, v2, v3, u, ua, u1, u2, u3, u4;
  Subset all;
  layoutSubset(&all, &layout, "all");
  fieldNew(&v1, &layout, nelemV);
  fieldNew(&v2, &layout, nelemV);
  fieldNew(&v3, &layout, nelemV);
  fieldNew(&u, &layout, nelemM);
  fieldNew(&ua, &layout, nelemM);
  fieldNew(&u1, &layout, nelemM);
  fieldNew(&u2, &layout, nelemM);
  fieldNew(&u3, &layout, nelemM);
  fieldNew(&u4, &layout, nelemM);

  double nv1 = 0, nu1 = 0;
  beginParallelSubset(all) {
    F_eq_r(&v1, 0);
    F_eq_r(&u, 0);
    F_eq_r(&u1, 0);
    for (int i = ti0; i < ti1; i++) {
      long kv = (long)i * nelemV * NI, km = (long)i * nelemM * NI;
      for (int j = 0; j < nelemV * NI; j++) v1.f[kv + j] = val(kv + j, 1);
      for (int j = 0; j < nelemM * NI; j++) {
        u.f[km + j] = val(km + j, 2);
        u1.f[km + j] = val(km + j, 3);
      }
    }
    P(makeU)(u.f, nc, ti0, ti1);
    F_eq_r(&u2, 0);
    double r[2];
    r[0] = lnorm2_F(&v1);
    r[1] = lnorm2_F(&u1);
    sumDoubleArray(r, 2);
    singleThread
    {
      nv1 = r[0];
      nu1 = r[1];
    }
  }
  endParallelSubset;
  printf0("|v1|^2 = %.6g  |u1|^2 = %.6g\n", nv1, nu1);

  /* makeU: U^+ U = 1 at every site and lane */
  {
    double worst = 0;
    for (long o = 0; o < layout.nSitesOuter; o++)
      for (int ii = 0; ii < NI; ii++) {
        const real *m = u.f + (long)nelemM * NI * o + ii;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            double re = 0, im = 0;
            for (int k = 0; k < 3; k++) {
              double ar = m[NI * 2 * (k + 3 * a)], ai = m[NI * (2 * (k + 3 * a) + 1)];
              double br = m[NI * 2 * (k + 3 * b)], bi = m[NI * (2 * (k + 3 * b) + 1)];
              re += ar * br + ai * bi;
              im += ar * bi - ai * br;
            }
            double d = fabs(re - (a == b)) + fabs(im);
            if (d > worst) worst = d;
          }
      }
    check("makeU: max |U^+U - 1|", worst, 8 * TOL);
  }

  int dir = 3;
  gaugeDoublestore((Field *[1]){&ua}, (Field *[1]){&u}, (int[1]){dir}, (int[1]){-1}, 1);

  Transporter T0, M0, Tb0;
  ShiftIndices *si0 = P(getShift)(dir, 1, "all");
  ShiftIndices *sib0 = P(getShift)(dir, -1, "all");
  transporterNew(&T0, u.f, si0, &all, nelemV, nc);
  transporterNew(&Tb0, ua.f, sib0, &all, nelemV, nc);
  transporterNew(&M0, u.f, si0, &all, nelemM, nc);

  /* 1. forward, then back with the doublestored adjoint */
  beginParallelSubset(all) {
    F_eq_r(&v2, 0);
    F_eq_r(&v3, 0);
    F_eq_r(&u2, 0);
  }
  endParallelSubset;
  transporterDo(&T0, v2.f, v1.f, 2);
  {
    real *sh = myalloc(sizeof(real) * nelemV * vol), *want = myalloc(sizeof(real) * nelemV * vol);
    qopShift(&layout, sh, v1.f, nelemV, dir + 1, 1);   /* sh(x) = v1(x + dir) */
    mat_vec(want, u.f, sh);
    check("transport vec td=2 vs U(x) v1(x+dir)", reldiff(v2.f, want, (long)nelemV * vol), TOL);
    free(sh);
    free(want);
  }
  transporterDo(&Tb0, v3.f, v2.f, 2);
  double nd = 0;
  beginParallelSubset(all) {
    X_vpeq_r_times_X(v3.f, -1, v1.f, v1.nelem, ti0, ti1);
    double r = lnorm2_F(&v3);
    sumDoubleArray(&r, 1);
    singleThread { nd = r; }
  }
  endParallelSubset;
  check("forward then backward returns v1", sqrt(nd / nv1), TOL);

  /* 2. matrix transport td=2 vs qopTransportM; td=0 vs qopShift */
  transporterDo(&M0, u2.f, u1.f, 2);
  qopTransportM(&layout, u4.f, u.f, u1.f, dir, 1);
  check("transport mat td=2 vs qopTransportM", reldiff(u2.f, u4.f, (long)nelemM * vol), TOL);
  transporterDo(&M0, u3.f, u1.f, 0);
  qopShift(&layout, u4.f, u1.f, nelemM, dir + 1, 1);
  check("transport mat td=0 vs qopShift", reldiff(u3.f, u4.f, (long)nelemM * vol), 0);

  /* 3. split-phase transport inside a team, even subset only */
  {
    Subset even;
    layoutSubset(&even, &layout, "even");
    Transporter Te;
    transporterNew(&Te, u.f, si0, &even, nelemV, nc);
    memset(v3.f, 0, sizeof(real) * nelemV * vol);
    beginParallel
    {
      transporterStartT(&Te, v3.f, v1.f, 2, TC);
      transporterLocalT(&Te, TC);
      transporterWaitT(&Te, TC);
    }
    endParallel;
    long ne = (long)nelemV * NI * layout.nEvenOuter;
    double odd = 0;
    for (long i = ne; i < (long)nelemV * vol; i++) odd += fabs(v3.f[i]);
    check("StartT/LocalT/WaitT even sites", reldiff(v3.f, v2.f, ne), TOL);
    check("StartT/LocalT/WaitT odd sites untouched", odd, 0);
    transporterFree(&Te);
  }

  /* 4. makeShift: arbitrary displacement, dst(x) = src(x - disp) */
  {
    int disp[4] = {1, -2, 0, 3};
    ShiftIndices sim;
    makeShift(&sim, &layout, disp);
    Transporter Tm, Tmu;
    transporterNew(&Tm, NULL, &sim, &all, nelemV, nc);
    transporterNew(&Tmu, u.f, &sim, &all, nelemV, nc);
    real *t1 = myalloc(sizeof(real) * nelemV * vol), *t2 = myalloc(sizeof(real) * nelemV * vol);
    /* x - disp = x - e0 + 2 e1 - 3 e3, composed from qopShift (x(s) = y(s + len*dir)) */
    qopShift(&layout, t1, v1.f, nelemV, 1, -1);
    qopShift(&layout, t2, t1, nelemV, 2, 2);
    qopShift(&layout, t1, t2, nelemV, 4, -3);
    transporterDo(&Tm, v3.f, v1.f, 0);
    check("makeShift transport td=0 vs qopShift chain", reldiff(v3.f, t1, (long)nelemV * vol), 0);
    mat_vec(t2, u.f, t1);
    for (long i = 0; i < (long)nelemV * vol; i++) t2[i] += v2.f[i];
    memcpy(v3.f, v2.f, sizeof(real) * nelemV * vol);
    transporterDo(&Tmu, v3.f, v1.f, 2);
    check("makeShift transport td=2 accumulates U src", reldiff(v3.f, t2, (long)nelemV * vol), TOL);
    int bad = 0;
    for (int o = 0; o < layout.nSitesOuter; o++) {
      int x[4];
      LayoutIndex li = {0, o * NI}, ls2;
      layoutCoord(&layout, x, &li);
      for (int k = 0; k < 4; k++) x[k] = ((x[k] - disp[k]) % ls[k] + ls[k]) % ls[k];
      layoutIndex(&layout, &ls2, x);
      bad += sim.sidx[o] != ls2.index / NI;
      bad += (sim.pidx[o] >= 0 ? sim.pidx[o] : -sim.pidx[o] - 2) != ls2.index / NI;
    }
    check("makeShift site table", bad, 0);
    ShiftIndices sie;
    makeShiftSub(&sie, &layout, disp, "even");
    int outside = 0;
    for (int o = layout.nEvenOuter; o < layout.nSitesOuter; o++) outside += sie.sidx[o] != -1;
    check("makeShiftSub even: odd sites unset", outside, 0);
    freeShift(&sie);
    transporterFree(&Tm);
    transporterFree(&Tmu);
    freeShift(&sim);
    free(t1);
    free(t2);
  }

  /* 5. host Linop (setOffSubC2 + gsum) in cgSolve vs solve2 */
  {
    Field U[4], Ub[4];
    real *uu[8];
    for (int mu = 0; mu < 4; mu++) {
      fieldNew(&U[mu], &layout, nelemM);
      fieldNew(&Ub[mu], &layout, nelemM);
      for (long i = 0; i < (long)nelemM * vol; i++) U[mu].f[i] = val(i, 10 + mu);
    }
    beginParallelSubset(all) {
      for (int mu = 0; mu < 4; mu++) P(makeU)(U[mu].f, nc, ti0, ti1);
    }
    endParallelSubset;
    Field *up[4] = {&U[0], &U[1], &U[2], &U[3]}, *ubp[4] = {&Ub[0], &Ub[1], &Ub[2], &Ub[3]};
    gaugeDoublestore(ubp, up, (int[4]){0, 1, 2, 3}, (int[4]){-1, -1, -1, -1}, 4);
    for (int mu = 0; mu < 4; mu++) {
      for (long i = 0; i < (long)nelemM * vol; i++) Ub[mu].f[i] = -Ub[mu].f[i];   /* -U(x-mu)^+ */
      uu[2 * mu] = U[mu].f;
      uu[2 * mu + 1] = Ub[mu].f;
    }
    double mass = 0.4;
    P(StaggeredLinks) sle, slo;
    P(createStaggeredLinks)(&sle, &layout, uu, NULL, "even");
    P(createStaggeredLinks)(&slo, &layout, uu, NULL, "odd");
    real *b = myalloc(sizeof(real) * nelemV * vol), *x1 = myalloc(sizeof(real) * nelemV * vol);
    real *x2 = myalloc(sizeof(real) * nelemV * vol), *tmp = myalloc(sizeof(real) * nelemV * vol);
    for (long i = 0; i < (long)nelemV * vol; i++) {
      b[i] = val(i, 20);
      x1[i] = x2[i] = tmp[i] = 0;
    }
    P(StaggeredSolver) ss;
    P(createStaggeredSolver)(&ss, &layout, uu, NULL, "even");
    ss.noscale = 1;
    int its2 = P(solve2)(&ss, x2, mass, b, SOLVE_RSQ, 2000, "even");
    LinArgs la = {&sle, &slo, tmp, mass * mass, {0}, {0}};
    layoutSubset(&la.sube, &layout, "even");
    layoutSubset(&la.subo, &layout, "odd");
    P(Cgls) cg;
    P(cglsInit)(&cg, nelemV * layout.nEvenOuter);
    int its1 = P(cgSolve)(&cg, x1, b, normalLinop, &la, NULL, NULL, SOLVE_RSQ, 2000);
    long ne = (long)nelemV * NI * layout.nEvenOuter;
    printf0("cgSolve with host Linop: %d its, solve2: %d its\n", its1, its2);
    check("Linop (setOffSubC2, gsum) cgSolve vs solve2", reldiff(x1, x2, ne), TOL_SOLVE);
    check("Linop cgSolve iterations vs solve2", abs(its1 - its2) > 1, 0);
    P(cglsFree)(&cg);
    P(freeStaggeredSolver)(&ss);
    P(freeStaggeredLinks)(&sle);
    P(freeStaggeredLinks)(&slo);
    free(b);
    free(x1);
    free(x2);
    free(tmp);
    for (int mu = 0; mu < 4; mu++) {
      fieldFree(&U[mu]);
      fieldFree(&Ub[mu]);
    }
  }

  transporterFree(&T0);
  transporterFree(&Tb0);
  transporterFree(&M0);
  fieldFree(&v1);
  fieldFree(&v2);
  fieldFree(&v3);
  fieldFree(&u);
  fieldFree(&ua);
  fieldFree(&u1);
  fieldFree(&u2);
  fieldFree(&u3);
  fieldFree(&u4);
  printf0("%s (%d failures)\n", nfail ? "FAILED" : "PASSED", nfail);
  return nfail;
}
