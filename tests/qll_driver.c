/*
 * Drop-in check of the QLL C API (qll.h) served by libqll_amd, written the
 * way QLL's own testdslash.c / testwilson.c drive it (testdslash.c:159-441):
 * stagDslashSetup, shifts, doublestored links with the negative adjoint,
 * smearFat7, dslash2 on all/even/odd for doublestored and Fb links, solve2,
 * solveMulti2, cgSolve with a host Linop, Wilson Dslash and solve.  Where
 * QLL compares with QOPQDP, this driver compares with the FP64 oracle
 * (oracle/oracle.c) in QLL's host layout.
 *
 * usage: qll_driver <D|F> <vlen> <Lx> <Ly> <Lz> <Lt>
 * prints one line per check ("ok"/"FAIL name value") and exits non-zero on
 * any failure.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "qll.h"

static int nfail = 0;
static void check(const char *name, double v, double tol)
{
  int ok = (v <= tol) && !isnan(v);
  printf("%s %-34s %.3e (tol %.1e)\n", ok ? "ok  " : "FAIL", name, v, tol);
  if (!ok) nfail++;
}

static OrcLayout olay;
static Layout layout;
static int dims[4];
static long V;

/* QLL host array <-> oracle natural layout */
static void nat2qll_d(const double *nat, double *q, int nelem) { orc_to_qll(&olay, nat, q, nelem); }
static void qll2nat_d(const double *q, double *nat, int nelem) { orc_from_qll(&olay, q, nat, nelem); }

static double relerr_sub(const double *a, const double *b, int nelem, int par)
{
  double num = 0, den = 0;
  for (long s = 0; s < V; s++) {
    if (par >= 0 && orc_parity(dims, s) != par) continue;
    for (int k = 0; k < nelem; k++) {
      double d = a[s * nelem + k] - b[s * nelem + k];
      num += d * d;
      den += b[s * nelem + k] * b[s * nelem + k];
    }
  }
  return sqrt(num / (den > 0 ? den : 1));
}

/* a diagonal host Linop in QLL layout, A = diag(w), with Anorm2 = p.Ap as
 * QLL's operators provide it */
typedef struct {
  double *w;
  long n;
} DiagArgs;

#define DEFINE_DIAG(R, SUF)                                                                    \
  static void diagOp##SUF(R *Ax, R *x, LinopInfo *li, void *args, ThreadCtx *TC)              \
  {                                                                                            \
    DiagArgs *a = args;                                                                        \
    long i0 = (TC->tid * a->n) / TC->nid, i1 = ((TC->tid + 1) * a->n) / TC->nid;              \
    double s = 0;                                                                              \
    for (long i = i0; i < i1; i++) {                                                           \
      Ax[i] = (R)(a->w[i] * x[i]);                                                             \
      s += (double)x[i] * Ax[i];                                                               \
    }                                                                                          \
    gsum(&s, 1, TC);                                                                           \
    if (li->wantAnorm2) li->Anorm2 = s;                                                        \
  }
DEFINE_DIAG(double, D)
DEFINE_DIAG(float, F)

#define PD(x) x##D
#define PF(x) x##F
#define PRECISION_BODY(R, P, TOL, DIAG)                                                        \
  {                                                                                            \
    const double tol = TOL;                                                                    \
    P(stagDslashSetup)(&layout, 4, dims, NULL);                                                \
    for (int i = 0; i < 4; i++) {                                                              \
      olay.innerGeom[i] = layout.innerGeom[i];                                                 \
    }                                                                                          \
    if (orc_layout_setup(&olay, dims, layout.innerGeom)) {                                     \
      printf("FAIL oracle layout\n");                                                          \
      return 1;                                                                                \
    }                                                                                          \
    check("layout nSitesInner", fabs(olay.nSitesInner - layout.nSitesInner), 0);               \
    /* links: seeded SU(3) + phases + fat7 + Naik, doublestored with -U^+ */                   \
    double *u = malloc(4 * V * 18 * 8), *fl = malloc(4 * V * 18 * 8), *ll = malloc(4 * V * 18 * 8); \
    double *l16 = malloc(16 * V * 18 * 8);                                                     \
    orc_random_su3_field(dims, 5, u);                                                          \
    orc_apply_stag_phases(dims, u);                                                            \
    double coef[6] = {5.0 / 8, -1.0 / 16, 1.0 / 64, -1.0 / 384, 0, -1.0 / 24};                \
    orc_fat7(dims, coef, u, u, fl, ll);                                                        \
    orc_stag_doublestore(dims, fl, ll, l16);                                                   \
    R *uq[8], *u3q[8];                                                                         \
    double *tmpq = malloc(V * 18 * 8), *tmpn = malloc(V * 24 * 8);                             \
    for (int k = 0; k < 8; k++) {                                                              \
      uq[k] = myalloc(V * 18 * sizeof(R));                                                     \
      u3q[k] = myalloc(V * 18 * sizeof(R));                                                    \
      nat2qll_d(l16 + (2 * k) * V * 18, tmpq, 18);                                             \
      for (long i = 0; i < V * 18; i++) uq[k][i] = (R)tmpq[i];                                 \
      nat2qll_d(l16 + (2 * k + 1) * V * 18, tmpq, 18);                                         \
      for (long i = 0; i < V * 18; i++) u3q[k][i] = (R)tmpq[i];                                \
    }                                                                                          \
    /* shift: backward link = shift of forward, as testdslash.c:255-285 builds it */           \
    {                                                                                          \
      R *sh = myalloc(V * 18 * sizeof(R));                                                     \
      double *thin = malloc(V * 18 * 8), *shn = malloc(V * 18 * 8), *ref = malloc(V * 18 * 8); \
      R *thq = myalloc(V * 18 * sizeof(R));                                                    \
      memcpy(thin, fl + 1 * V * 18, V * 18 * 8);                                               \
      nat2qll_d(thin, tmpq, 18);                                                               \
      for (long i = 0; i < V * 18; i++) thq[i] = (R)tmpq[i];                                   \
      for (int len = 1; len <= 3; len += 2)                                                    \
        for (int sg = -1; sg <= 1; sg += 2) {                                                  \
          P(shift)(sh, thq, 18, sg * 2, len);                                                  \
          for (long i = 0; i < V * 18; i++) tmpq[i] = sh[i];                                   \
          qll2nat_d(tmpq, shn, 18);                                                            \
          for (long s = 0; s < V; s++) {                                                       \
            long n = orc_neighbor(dims, s, 1, sg * len);                                       \
            for (int e = 0; e < 18; e++) ref[s * 18 + e] = (R)thin[n * 18 + e];                \
          }                                                                                    \
          check(sg > 0 ? "shift +y" : "shift -y", relerr_sub(shn, ref, 18, -1), 0);           \
        }                                                                                      \
      free(thin); free(shn); free(ref); free(sh); free(thq);                                   \
    }                                                                                          \
    /* smearFat7 through the API vs the oracle */                                              \
    {                                                                                          \
      P(Field) fu[4], ff[4], fL[4], *fup[4], *ffp[4], *fLp[4];                                 \
      double *un = malloc(4 * V * 18 * 8);                                                     \
      for (int mu = 0; mu < 4; mu++) {                                                         \
        P(fieldNew)(&fu[mu], &layout, 18);                                                     \
        P(fieldNew)(&ff[mu], &layout, 18);                                                     \
        P(fieldNew)(&fL[mu], &layout, 18);                                                     \
        fup[mu] = &fu[mu]; ffp[mu] = &ff[mu]; fLp[mu] = &fL[mu];                               \
        nat2qll_d(u + mu * V * 18, tmpq, 18);                                                  \
        for (long i = 0; i < V * 18; i++) fu[mu].f[i] = (R)tmpq[i];                            \
      }                                                                                        \
      Fat7Coeffs c7 = {coef[0], coef[1], coef[2], coef[3], coef[4], coef[5]};                  \
      P(smearFat7)(ffp, fLp, &c7, fup, fup);                                                   \
      double ef = 0, el = 0;                                                                   \
      for (int mu = 0; mu < 4; mu++) {                                                         \
        for (long i = 0; i < V * 18; i++) tmpq[i] = ff[mu].f[i];                               \
        qll2nat_d(tmpq, un, 18);                                                               \
        double e1 = relerr_sub(un, fl + mu * V * 18, 18, -1);                                  \
        for (long i = 0; i < V * 18; i++) tmpq[i] = fL[mu].f[i];                               \
        qll2nat_d(tmpq, un, 18);                                                               \
        double e2 = relerr_sub(un, ll + mu * V * 18, 18, -1);                                  \
        if (e1 > ef) ef = e1;                                                                  \
        if (e2 > el) el = e2;                                                                  \
      }                                                                                        \
      check("smearFat7 fat", ef, tol);                                                         \
      check("smearFat7 long", el, tol);                                                        \
      double pl[6], plr[6];                                                                    \
      P(gaugePlaq)(pl, fup);                                                                   \
      orc_plaquette(dims, u, plr);                                                             \
      double pe = 0;                                                                           \
      for (int i = 0; i < 6; i++) pe = fmax(pe, fabs(pl[i] - plr[i]));                          \
      check("gaugePlaq", pe, tol);                                                             \
      /* gaugeDoublestore: ua = (u shifted from x-mu)^+ , compared with the oracle's +U^+ */   \
      double *w8 = malloc(8 * V * 18 * 8);                                                     \
      orc_wilson_doublestore(dims, u, w8);                                                     \
      int dirs[4] = {0, 1, 2, 3}, lens[4] = {-1, -1, -1, -1};                                  \
      P(gaugeDoublestore)(ffp, fup, dirs, lens, 4);                                            \
      double ed = 0;                                                                           \
      for (int mu = 0; mu < 4; mu++) {                                                         \
        for (long i = 0; i < V * 18; i++) tmpq[i] = ff[mu].f[i];                               \
        qll2nat_d(tmpq, un, 18);                                                               \
        ed = fmax(ed, relerr_sub(un, w8 + (2 * mu + 1) * V * 18, 18, -1));                     \
      }                                                                                        \
      check("gaugeDoublestore", ed, tol);                                                      \
      /* field utilities */                                                                    \
      Subset all, ev;                                                                          \
      layoutSubset(&all, &layout, "all");                                                      \
      layoutSubset(&ev, &layout, "even");                                                      \
      double n2 = P(fieldNorm2)(&fu[0], &all), n2r = 0;                                        \
      for (long i = 0; i < V * 18; i++) n2r += (double)fu[0].f[i] * fu[0].f[i];                \
      check("fieldNorm2", fabs(n2 - n2r) / n2r, 1e-13);                                        \
      P(fieldZero)(&fu[0], &ev);                                                               \
      double n2e = P(fieldNorm2)(&fu[0], &ev);                                                 \
      check("fieldZero even", n2e, 0);                                                         \
      for (int mu = 0; mu < 4; mu++) {                                                         \
        P(fieldFree)(&fu[mu]); P(fieldFree)(&ff[mu]); P(fieldFree)(&fL[mu]);                   \
      }                                                                                        \
      free(un); free(w8);                                                                      \
    }                                                                                          \
    /* dslash2 on all/even/odd, doublestored and Fb links (testdslash.c:122-156,330-351) */    \
    double *yn = malloc(V * 6 * 8), *xn = malloc(V * 6 * 8), *rn = malloc(V * 6 * 8);          \
    R *yq = myalloc(V * 6 * sizeof(R)), *xq = myalloc(V * 6 * sizeof(R));                      \
    orc_gauss_vector(dims, 9, 100, -1, 6, yn);                                                 \
    nat2qll_d(yn, tmpq, 6);                                                                    \
    for (long i = 0; i < V * 6; i++) yq[i] = (R)tmpq[i];                                       \
    for (long i = 0; i < V * 6; i++) tmpq[i] = yq[i];                                          \
    qll2nat_d(tmpq, yn, 6);                                                                    \
    const char *subs[3] = {"all", "even", "odd"};                                              \
    for (int fb = 0; fb < 2; fb++)                                                             \
      for (int si = 0; si < 3; si++) {                                                         \
        P(StaggeredLinks) sl;                                                                  \
        if (fb) P(createStaggeredLinksFb)(&sl, &layout, uq, u3q, (char *)subs[si]);            \
        else P(createStaggeredLinks)(&sl, &layout, uq, u3q, (char *)subs[si]);                 \
        double mass = 0.5;                                                                     \
        for (long i = 0; i < V * 6; i++) xq[i] = (R)1234.5;                                    \
        P(dslash2)(&sl, xq, mass, yq);                                                         \
        for (long i = 0; i < V * 6; i++) tmpq[i] = xq[i];                                      \
        qll2nat_d(tmpq, xn, 6);                                                                \
        for (long i = 0; i < V * 6; i++) rn[i] = 1234.5;                                       \
        orc_dslash(dims, l16, 16, mass, yn, rn, si);                                           \
        char nm[64];                                                                           \
        snprintf(nm, sizeof nm, "dslash2 %s%s", subs[si], fb ? " Fb" : "");                    \
        check(nm, relerr_sub(xn, rn, 6, -1), fb ? 10 * tol : tol);                             \
        P(freeStaggeredLinks)(&sl);                                                            \
      }                                                                                        \
    /* solve2 and solveMulti2 (testdslash.c:355-437) */                                        \
    {                                                                                          \
      P(StaggeredSolver) ss;                                                                   \
      P(createStaggeredSolver)(&ss, &layout, uq, u3q, "even");                                 \
      double mass = 0.2, rsq = sizeof(R) == 8 ? 1e-16 : 1e-10;                                 \
      double *bn = malloc(V * 6 * 8), *xr = malloc(V * 6 * 8);                                 \
      orc_gauss_vector(dims, 4, 100, 0, 6, bn);                                                \
      nat2qll_d(bn, tmpq, 6);                                                                  \
      for (long i = 0; i < V * 6; i++) { yq[i] = (R)tmpq[i]; }                                 \
      for (long i = 0; i < V * 6; i++) tmpq[i] = yq[i];                                        \
      qll2nat_d(tmpq, bn, 6);                                                                  \
      int its = P(solve2)(&ss, xq, mass, yq, rsq, 2000, "even");                               \
      double fr;                                                                               \
      int itr = orc_cg(dims, l16, 16, mass, bn, xr, rsq, 2000, 0, &fr);                        \
      check("solve2 iterations", fabs(its - itr), sizeof(R) == 8 ? 1 : 0.05 * itr + 2);        \
      check("solve2 ss->its", fabs(ss.its - its), 0);                                          \
      for (long i = 0; i < V * 6; i++) tmpq[i] = xq[i];                                        \
      qll2nat_d(tmpq, xn, 6);                                                                  \
      check("solve2 solution", relerr_sub(xn, xr, 6, 0), sizeof(R) == 8 ? 1e-6 : 1e-3);        \
      double masses[5] = {0.2, 0.28, 0.392, 0.5488, 0.76832};                                 \
      R *xs[5];                                                                                \
      double *xrs[5];                                                                          \
      for (int i = 0; i < 5; i++) { xs[i] = myalloc(V * 6 * sizeof(R)); xrs[i] = malloc(V * 6 * 8); } \
      int itm = P(solveMulti2)(&ss, xs, masses, 5, yq, rsq, 2000, "even");                      \
      int itmr = orc_cgms(dims, l16, 16, masses, 5, bn, xrs, rsq, 2000, 0);                    \
      check("solveMulti2 iterations", fabs(itm - itmr), sizeof(R) == 8 ? 1 : 0.05 * itmr + 2); \
      double em = 0;                                                                           \
      for (int i = 0; i < 5; i++) {                                                            \
        for (long k = 0; k < V * 6; k++) tmpq[k] = xs[i][k];                                   \
        qll2nat_d(tmpq, xn, 6);                                                                \
        em = fmax(em, relerr_sub(xn, xrs[i], 6, 0));                                           \
        free(xs[i]); free(xrs[i]);                                                             \
      }                                                                                        \
      check("solveMulti2 solutions", em, sizeof(R) == 8 ? 1e-6 : 1e-3);                        \
      /* noscale: x without the factor mass */                                                 \
      ss.noscale = 1;                                                                          \
      P(solve2)(&ss, xq, mass, yq, rsq, 2000, "even");                                         \
      for (long i = 0; i < V * 6; i++) tmpq[i] = xq[i];                                        \
      qll2nat_d(tmpq, xn, 6);                                                                  \
      for (long i = 0; i < V * 6; i++) xr[i] /= mass;                                          \
      check("solve2 noscale", relerr_sub(xn, xr, 6, 0), sizeof(R) == 8 ? 1e-6 : 1e-3);         \
      P(freeStaggeredSolver)(&ss);                                                             \
      free(bn); free(xr);                                                                      \
    }                                                                                          \
    /* cgSolve / cgmsSolve with a host Linop (lib/cg.c:124-136, lib/cgms.c:181-195) */         \
    {                                                                                          \
      int nelem = 6 * layout.nEvenOuter;                                                       \
      long n = (long)nelem * layout.nSitesInner;                                               \
      DiagArgs da = {malloc(n * 8), n};                                                        \
      R *b = myalloc(n * sizeof(R)), *x = myalloc(n * sizeof(R));                              \
      for (long i = 0; i < n; i++) { da.w[i] = 1.0 + (i % 7) * 0.5; b[i] = (R)(1.0 + 0.01 * (i % 13)); } \
      P(Cgls) cg;                                                                              \
      P(cglsInit)(&cg, nelem);                                                                 \
      int it = P(cgSolve)(&cg, x, b, DIAG, &da, NULL, NULL, 1e-20, 200);                       \
      double e = 0;                                                                            \
      for (long i = 0; i < n; i++) e = fmax(e, fabs(x[i] - b[i] / da.w[i]));                   \
      check("cgSolve host Linop", e, sizeof(R) == 8 ? 1e-12 : 1e-5);                           \
      check("cgSolve its (7 distinct eigenvalues)", it > (sizeof(R) == 8 ? 8 : 10) ? it : 0, 0);                  \
      double sh[3] = {0, 0.5, 2.0}, rs[3] = {1e-20, 1e-20, 1e-20};                             \
      R *xm[3] = {myalloc(n * sizeof(R)), myalloc(n * sizeof(R)), myalloc(n * sizeof(R))};    \
      P(cgmsSolve)(&cg, xm, b, sh, 3, DIAG, &da, NULL, NULL, rs, 200);                          \
      double e2 = 0;                                                                           \
      for (int k = 0; k < 3; k++)                                                              \
        for (long i = 0; i < n; i++) e2 = fmax(e2, fabs(xm[k][i] - b[i] / (da.w[i] + sh[k]))); \
      check("cgmsSolve host Linop", e2, sizeof(R) == 8 ? 1e-11 : 1e-5);                        \
      P(cglsFree)(&cg);                                                                        \
      free(da.w); free(b); free(x); free(xm[0]); free(xm[1]); free(xm[2]);                     \
    }                                                                                          \
    /* Wilson (testwilson.c:257-276; links u[2mu+1] = +U(x-mu)^+) */                          \
    {                                                                                          \
      double *w8 = malloc(8 * V * 18 * 8);                                                     \
      double *un = malloc(4 * V * 18 * 8);                                                     \
      orc_random_su3_field(dims, 21, un);                                                      \
      orc_wilson_doublestore(dims, un, w8);                                                    \
      R *wq[8];                                                                                \
      for (int k = 0; k < 8; k++) {                                                            \
        wq[k] = myalloc(V * 18 * sizeof(R));                                                   \
        nat2qll_d(w8 + k * V * 18, tmpq, 18);                                                  \
        for (long i = 0; i < V * 18; i++) wq[k][i] = (R)tmpq[i];                               \
      }                                                                                        \
      double *sy = malloc(V * 24 * 8), *sx = malloc(V * 24 * 8), *sr = malloc(V * 24 * 8);     \
      R *syq = myalloc(V * 24 * sizeof(R)), *sxq = myalloc(V * 24 * sizeof(R));                \
      double *t24 = malloc(V * 24 * 8);                                                        \
      orc_gauss_vector(dims, 3, 200, -1, 24, sy);                                              \
      orc_to_qll(&olay, sy, t24, 24);                                                          \
      for (long i = 0; i < V * 24; i++) syq[i] = (R)t24[i];                                    \
      for (long i = 0; i < V * 24; i++) t24[i] = syq[i];                                       \
      orc_from_qll(&olay, t24, sy, 24);                                                        \
      for (int si = 0; si < 3; si++) {                                                         \
        P(WilsonLinks) wl;                                                                     \
        P(createWilsonLinks)(&wl, &layout, wq, (char *)subs[si]);                              \
        for (long i = 0; i < V * 24; i++) sxq[i] = (R)7.0;                                     \
        P(wilsonDslash)(&wl, sxq, 0.1, 1, syq);                                                \
        for (long i = 0; i < V * 24; i++) t24[i] = sxq[i];                                     \
        orc_from_qll(&olay, t24, sx, 24);                                                      \
        for (long i = 0; i < V * 24; i++) sr[i] = 7.0;                                         \
        orc_wilson_dslash(dims, w8, 0.1, sy, sr, si);                                          \
        char nm[64];                                                                           \
        snprintf(nm, sizeof nm, "wilsonDslash %s", subs[si]);                                  \
        check(nm, relerr_sub(sx, sr, 24, -1), tol);                                            \
        P(freeWilsonLinks)(&wl);                                                               \
      }                                                                                        \
      /* wilsonSolve: (m^2 - D_eo D_oe) x = y, x <- mass x; check (m + D) x_full = y */        \
      P(WilsonSolver) ws;                                                                      \
      P(createWilsonSolver)(&ws, &layout, wq, "even");                                         \
      double mass = 4.5, rsq = sizeof(R) == 8 ? 1e-20 : 1e-10;                                 \
      for (long s = 0; s < V; s++) if (orc_parity(dims, s)) for (int k = 0; k < 24; k++) sy[s * 24 + k] = 0; \
      orc_to_qll(&olay, sy, t24, 24);                                                          \
      for (long i = 0; i < V * 24; i++) syq[i] = (R)t24[i];                                    \
      for (long i = 0; i < V * 24; i++) sxq[i] = 0;                                            \
      P(wilsonSolve)(&ws, sxq, mass, syq, rsq, 5000, "even");                                  \
      for (long i = 0; i < V * 24; i++) t24[i] = sxq[i];                                       \
      orc_from_qll(&olay, t24, sx, 24);                                                        \
      /* x_o = -D_oe x_e / m; then (m + D) x = y on all sites with D the massless Dslash */    \
      double *xo = calloc(V * 24, 8), *full = calloc(V * 24, 8);                               \
      orc_wilson_dslash(dims, w8, -4.0, sx, xo, 2);                                            \
      for (long s = 0; s < V; s++)                                                             \
        for (int k = 0; k < 24; k++)                                                           \
          full[s * 24 + k] = orc_parity(dims, s) ? -xo[s * 24 + k] / mass : sx[s * 24 + k];    \
      double *res = calloc(V * 24, 8);                                                         \
      orc_wilson_dslash(dims, w8, -4.0, full, res, 0);                                         \
      for (long i = 0; i < V * 24; i++) res[i] += mass * full[i];                              \
      check("wilsonSolve (m+D)x = y", relerr_sub(res, sy, 24, -1), sizeof(R) == 8 ? 1e-9 : 1e-4); \
      P(freeWilsonSolver)(&ws);                                                                \
      free(xo); free(full); free(res); free(sy); free(sx); free(sr); free(t24); free(w8); free(un); \
      for (int k = 0; k < 8; k++) free(wq[k]);                                                 \
      free(syq); free(sxq);                                                                    \
    }                                                                                          \
    for (int k = 0; k < 8; k++) { free(uq[k]); free(u3q[k]); }                                 \
    free(u); free(fl); free(ll); free(l16); free(tmpq); free(tmpn);                            \
    free(yn); free(xn); free(rn); free(yq); free(xq);                                          \
  }


int main(int argc, char **argv)
{
  if (argc < 7) {
    fprintf(stderr, "usage: %s <D|F> <vlen> Lx Ly Lz Lt\n", argv[0]);
    return 2;
  }
  int vlen = atoi(argv[2]);
  for (int i = 0; i < 4; i++) dims[i] = atoi(argv[3 + i]);
  V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  layout.nranks = 1;
  layout.myrank = 0;
  layout.nDim = 4;
  qllSetVlen(vlen, vlen);
  if (argv[1][0] == 'D') {
    PRECISION_BODY(double, PD, 1e-13, diagOpD)
  } else {
    PRECISION_BODY(float, PF, 1e-5, diagOpF)
  }
  printf("%s: %d failure(s)\n", nfail ? "FAILED" : "PASSED", nfail);
  return nfail ? 1 : 0;
}
