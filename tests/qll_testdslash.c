/*
 * A QLL test program in the shape of QLL's testdslash.c and testwilson.c
 * (testdslash.c:159-441, testwilson.c:140-366), linked against libqll_amd:
 * QLL's API computes, and the qop* entry points the reference's tests compare
 * with give the expected values.  Those are served here by the FP64 oracle in
 * QLL's layout (oracle/qop.c; QOPQDP is absent).
 *
 * Flow: stagDslashSetup; backward links by shift and negative adjoint,
 * checked against qopShift; smearFat7 vs qopSmear; dslash2 on all/even/odd
 * with doublestored and Fb links vs qopDslash; solve2 vs qopSolve;
 * solveMulti2 vs qopSolveMulti; the qop projector identities; Wilson Dslash
 * and solve vs qopWilsonDslash / qopWilsonSolve.
 *
 * Build with -DDOUBLE for the D API (link libqopD), without for F (libqopF).
 * usage: qll_testdslash <vlen> Lx Ly Lz Lt    (exit status: failures)
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "qop.h"

#ifdef DOUBLE
#define P(x) x##D
#define TOL_OP 1e-12
#define TOL_SOLVE 1e-6
#define SOLVE_RSQ 1e-20
#else
#define P(x) x##F
#define TOL_OP 2e-6
#define TOL_SOLVE 2e-3
#define SOLVE_RSQ 1e-11
#endif
typedef QOP_REAL R;

static Layout layout;
static int nfail;
static long vol;

/* relative rms difference of two whole arrays (QLL layout) */
static void compare(const char *what, const R *got, const R *want, long n, double tol)
{
  double d2 = 0, w2 = 0;
  for (long i = 0; i < n; i++) {
    double d = (double)got[i] - (double)want[i];
    d2 += d * d;
    w2 += (double)want[i] * want[i];
  }
  double e = sqrt(d2 / (w2 > 0 ? w2 : 1));
  int ok = e <= tol && !isnan(e);
  printf("%s %-36s %.3e (tol %.1e)\n", ok ? "ok  " : "FAIL", what, e, tol);
  nfail += !ok;
}

/* deterministic fill of the whole array, different per (tag, index) */
static void fill(R *v, long n, int tag)
{
  for (long i = 0; i < n; i++) v[i] = (R)(0.5 * sin(0.7 * (double)i + 1.3 * tag) + 0.2 * cos(0.11 * (double)i * tag));
}

/* x = -x^+ for every 3x3 matrix of a QLL-layout array (testdslash.c:60-85) */
static void neg_adj(R *x)
{
  int NI = layout.nSitesInner;
  for (int o = 0; o < layout.nSitesOuter; o++)
    for (int ii = 0; ii < NI; ii++) {
      R *m = x + (long)18 * NI * o + ii;
      double t[18];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {   /* element (r,c) at reals 2(3c+r) */
          t[2 * (3 * c + r)] = -(double)m[NI * 2 * (3 * r + c)];
          t[2 * (3 * c + r) + 1] = (double)m[NI * (2 * (3 * r + c) + 1)];
        }
      for (int e = 0; e < 18; e++) m[NI * e] = (R)t[e];
    }
}

static R *alloc_r(long n) { return myalloc(n * sizeof(R)); }

static void test_dslash(P(StaggeredLinks) *sl, R *u[8], R *u3[8], R mass, R *y, char *sub, const char *kind)
{
  R *x = alloc_r(6 * vol), *t = alloc_r(6 * vol);
  fill(x, 6 * vol, 7);   /* both start from the same contents */
  memcpy(t, x, 6 * vol * sizeof(R));
  P(dslash2)(sl, x, mass, y);
  qopDslash(&layout, t, u, u3, mass, y, sub);
  char what[64];
  snprintf(what, sizeof what, "dslash2 %s %s", kind, sub);
  compare(what, x, t, 6 * vol, TOL_OP);
  free(x);
  free(t);
}

/* recon(proj(psi)) = (1 + s g_dir) psi (dir < 4), (1 + s g5)/2 psi (dir 4)
 * (lib/Gamma.lua:46-75), checked per site in QLL layout */
static void test_projectors(void)
{
  R *psi = alloc_r(24 * vol), *h = alloc_r(12 * vol), *rec = alloc_r(24 * vol), *want = alloc_r(24 * vol);
  fill(psi, 24 * vol, 11);
  int NI = layout.nSitesInner;
  for (int dir = 0; dir < 5; dir++)
    for (int sign = -1; sign <= 1; sign += 2) {
      double g[32];
      orc_gamma(dir, g);
      double f = dir == 4 ? 0.5 : 1.0;
      for (int o = 0; o < layout.nSitesOuter; o++)
        for (int ii = 0; ii < NI; ii++) {
          const R *p = psi + (long)24 * NI * o + ii;
          R *w = want + (long)24 * NI * o + ii;
          for (int r = 0; r < 4; r++)
            for (int c = 0; c < 3; c++) {
              double complex z = 0;
              for (int k = 0; k < 4; k++) {
                double complex t = (r == k ? 1.0 : 0.0) + sign * (g[2 * (4 * r + k)] + I * g[2 * (4 * r + k) + 1]);
                z += f * t * ((double)p[NI * (6 * k + 2 * c)] + I * (double)p[NI * (6 * k + 2 * c + 1)]);
              }
              w[NI * (6 * r + 2 * c)] = (R)creal(z);
              w[NI * (6 * r + 2 * c + 1)] = (R)cimag(z);
            }
        }
      qopSproj(&layout, h, psi, dir, sign);
      qopSrecon(&layout, rec, h, dir, sign);
      char what[64];
      snprintf(what, sizeof what, "qop srecon(sproj) dir %d sign %+d", dir, sign);
      compare(what, rec, want, 24 * vol, TOL_OP);
    }
  free(psi);
  free(h);
  free(rec);
  free(want);
}

int main(int argc, char **argv)
{
  if (argc < 6) {
    fprintf(stderr, "usage: %s <vlen> Lx Ly Lz Lt\n", argv[0]);
    return 2;
  }
  int vlen = atoi(argv[1]);
  int ls[4], rg[4];
  for (int i = 0; i < 4; i++) ls[i] = atoi(argv[2 + i]);
  qllSetVlen(vlen, vlen);
  qopStagDslashInit(&argc, &argv);
  layout.nranks = 1;
  layout.myrank = 0;
  layout.nDim = 4;
  layout.physGeom = ls;
  qopStagDslashSetup(&layout, rg);
  P(stagDslashSetup)(&layout, 4, ls, rg);
  vol = layout.nSites;
  printf("lattice %dx%dx%dx%d vlen %d (nSitesInner %d)\n", ls[0], ls[1], ls[2], ls[3], vlen, layout.nSitesInner);

  R *u[8], *u3[8], *tm = alloc_r(18 * vol), *v2 = alloc_r(6 * vol);
  for (int i = 0; i < 8; i++) {
    u[i] = alloc_r(18 * vol);
    u3[i] = alloc_r(18 * vol);
    fill(u[i], 18 * vol, 20 + i);
    fill(u3[i], 18 * vol, 40 + i);
  }
  fill(v2, 6 * vol, 3);

  /* backward links: shift of the forward ones, then the negative adjoint
   * (testdslash.c:255-285) */
  for (int mu = 0; mu < 4; mu++) {
    int f = 2 * mu;
    for (int sg = 1; sg >= -1; sg -= 2) {
      char what[64];
      P(shift)(u[f + 1], u[f], 18, sg * (mu + 1), 1);
      qopShift(&layout, tm, u[f], 18, sg * (mu + 1), 1);
      snprintf(what, sizeof what, "shift1 %+d", sg * (mu + 1));
      compare(what, u[f + 1], tm, 18 * vol, 0);
      P(shift)(u3[f + 1], u3[f], 18, sg * (mu + 1), 3);
      qopShift(&layout, tm, u3[f], 18, sg * (mu + 1), 3);
      snprintf(what, sizeof what, "shift3 %+d", sg * (mu + 1));
      compare(what, u3[f + 1], tm, 18 * vol, 0);
    }
    neg_adj(u[f + 1]);
    neg_adj(u3[f + 1]);
  }

  /* qopTransportM with unit matrices is a shift */
  {
    R *one = alloc_r(18 * vol), *a = alloc_r(18 * vol), *b = alloc_r(18 * vol);
    int NI = layout.nSitesInner;
    memset(one, 0, 18 * vol * sizeof(R));
    for (int o = 0; o < layout.nSitesOuter; o++)
      for (int ii = 0; ii < NI; ii++)
        for (int d = 0; d < 3; d++) one[(long)18 * NI * o + NI * 2 * (3 * d + d) + ii] = 1;
    for (int mu = 0; mu < 4; mu++) {
      qopTransportM(&layout, a, one, u[0], mu, 1);
      qopShift(&layout, b, u[0], 18, mu + 1, 1);
      compare("qopTransportM(1) fwd = shift", a, b, 18 * vol, 0);
      qopTransportM(&layout, a, one, u[0], mu, -1);
      qopShift(&layout, b, u[0], 18, -(mu + 1), 1);
      compare("qopTransportM(1) bwd = shift", a, b, 18 * vol, 0);
    }
    free(one);
    free(a);
    free(b);
  }

  /* fat7 with the Lepage coefficients of testdslash.c:300-315 */
  {
    P(Field) fl[4], ll[4], fl2[4], ll2[4], uu[4], *flp[4], *llp[4], *fl2p[4], *ll2p[4], *uup[4];
    for (int i = 0; i < 4; i++) {
      P(fieldNew)(&fl[i], &layout, 18);
      P(fieldNew)(&ll[i], &layout, 18);
      P(fieldNew)(&fl2[i], &layout, 18);
      P(fieldNew)(&ll2[i], &layout, 18);
      P(fieldSet)(&uu[i], u[2 * i], &layout, 18);
      flp[i] = &fl[i];
      llp[i] = &ll[i];
      fl2p[i] = &fl2[i];
      ll2p[i] = &ll2[i];
      uup[i] = &uu[i];
    }
    Fat7Coeffs coef = FAT7COEFFS_ZERO;
    coef.one_link = 0.3 - (0.1 * 15) / 16;
    coef.three_staple = 0.2 / (4 * 6);
    coef.five_staple = 0.2 / (16 * 24);
    coef.seven_staple = 0.2 / (64 * 48);
    coef.lepage = 0.1 / (16 * 6);
    coef.naik = 0.2;
    P(smearFat7)(flp, llp, &coef, uup, uup);
    qopSmear(fl2p, ll2p, &coef, uup, uup);
    for (int i = 0; i < 4; i++) {
      compare("smearFat7 fat", fl[i].f, fl2[i].f, 18 * vol, TOL_OP);
      compare("smearFat7 long", ll[i].f, ll2[i].f, 18 * vol, TOL_OP);
      P(fieldFree)(&fl[i]);
      P(fieldFree)(&ll[i]);
      P(fieldFree)(&fl2[i]);
      P(fieldFree)(&ll2[i]);
    }
  }

  /* Dslash on all/even/odd, doublestored and Fb links */
  R mass = 0.5;
  {
    char *subs[3] = {"all", "even", "odd"};
    for (int k = 0; k < 3; k++) {
      P(StaggeredLinks) sl, slfb;
      P(createStaggeredLinks)(&sl, &layout, u, u3, subs[k]);
      P(createStaggeredLinksFb)(&slfb, &layout, u, u3, subs[k]);
      test_dslash(&sl, u, u3, mass, v2, subs[k], "doublestored");
      test_dslash(&slfb, u, u3, mass, v2, subs[k], "Fb");
      P(freeStaggeredLinks)(&sl);
      P(freeStaggeredLinks)(&slfb);
    }
  }

  /* solve2 and solveMulti2 on even sites (testdslash.c:360-436) */
  {
    P(StaggeredSolver) sse;
    P(createStaggeredSolver)(&sse, &layout, u, u3, "even");
    R *v1 = alloc_r(6 * vol), *tv = alloc_r(6 * vol);
    fill(v1, 6 * vol, 5);
    memcpy(tv, v1, 6 * vol * sizeof(R));
    int its = P(solve2)(&sse, v1, mass, v2, SOLVE_RSQ, 5000, "even");
    printf("     solve2 iterations %d\n", its);
    qopSolve(&layout, tv, u, u3, mass, v2, SOLVE_RSQ, "even");
    compare("solve2 vs qopSolve", v1, tv, 6 * vol, TOL_SOLVE);
    enum { NM = 5 };
    double masses[NM];
    R *xs[NM], *ys[NM];
    for (int i = 0; i < NM; i++) {
      masses[i] = i ? 1.4 * masses[i - 1] : mass;
      xs[i] = alloc_r(6 * vol);
      ys[i] = alloc_r(6 * vol);
      fill(xs[i], 6 * vol, 60 + i);
      memcpy(ys[i], xs[i], 6 * vol * sizeof(R));
    }
    its = P(solveMulti2)(&sse, xs, masses, NM, v2, SOLVE_RSQ, 5000, "even");
    printf("     solveMulti2 iterations %d\n", its);
    qopSolveMulti(&layout, ys, u, u3, masses, v2, NM, SOLVE_RSQ, "even");
    for (int i = 0; i < NM; i++) {
      char what[64];
      snprintf(what, sizeof what, "solveMulti2 mass[%d] = %g", i, masses[i]);
      compare(what, xs[i], ys[i], 6 * vol, TOL_SOLVE);
      free(xs[i]);
      free(ys[i]);
    }
    P(freeStaggeredSolver)(&sse);
    free(v1);
    free(tv);
  }
  qopStagDslashFini();

  /* Wilson (testwilson.c:140-366): links U and +U^+ of the shifted forward
   * links, Dslash on all/even/odd, solve */
  qopWilsonDslashInit(&argc, &argv);
  qopWilsonDslashSetup(&layout);
  test_projectors();
  {
    R *w[8], *y = alloc_r(24 * vol), *x = alloc_r(24 * vol), *t = alloc_r(24 * vol);
    for (int i = 0; i < 8; i++) w[i] = alloc_r(18 * vol);
    for (int mu = 0; mu < 4; mu++) {
      memcpy(w[2 * mu], u[2 * mu], 18 * vol * sizeof(R));
      P(shift)(w[2 * mu + 1], w[2 * mu], 18, -(mu + 1), 1);
      neg_adj(w[2 * mu + 1]);   /* -(-U^+) below gives +U^+ */
      for (long i = 0; i < 18 * vol; i++) w[2 * mu + 1][i] = -w[2 * mu + 1][i];
    }
    fill(y, 24 * vol, 9);
    char *subs[3] = {"all", "even", "odd"};
    for (int k = 0; k < 3; k++) {
      P(WilsonLinks) wl;
      P(createWilsonLinks)(&wl, &layout, w, subs[k]);
      fill(x, 24 * vol, 10);
      memcpy(t, x, 24 * vol * sizeof(R));
      P(wilsonDslash)(&wl, x, mass, 1, y);
      qopWilsonDslash(&layout, t, w, mass, 1, y, subs[k]);
      char what[64];
      snprintf(what, sizeof what, "wilsonDslash %s", subs[k]);
      compare(what, x, t, 24 * vol, TOL_OP);
      P(freeWilsonLinks)(&wl);
    }
    P(WilsonSolver) wss;
    P(createWilsonSolver)(&wss, &layout, w, "even");
    R wmass = 10;   /* testwilson.c:296 */
    fill(x, 24 * vol, 12);
    memcpy(t, x, 24 * vol * sizeof(R));
    int its = P(wilsonSolve)(&wss, x, wmass, y, SOLVE_RSQ, 5000, "even");
    printf("     wilsonSolve iterations %d\n", its);
    qopWilsonSolve(&layout, t, w, wmass, y, SOLVE_RSQ, "even");
    compare("wilsonSolve vs qopWilsonSolve", x, t, 24 * vol, TOL_SOLVE);
    P(freeWilsonSolver)(&wss);
    for (int i = 0; i < 8; i++) free(w[i]);
    free(y);
    free(x);
    free(t);
  }
  qopWilsonDslashFini();

  for (int i = 0; i < 8; i++) {
    free(u[i]);
    free(u3[i]);
  }
  free(tm);
  free(v2);
  printf("%s: %d failure(s)\n", nfail ? "FAILED" : "PASSED", nfail);
  return nfail;
}
