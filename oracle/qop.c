/*
 * qop* entry points of QLL's test programs over the FP64 oracle (see qop.h).
 * TEST INFRASTRUCTURE ONLY: linked by tests/ programs, never by the product.
 *
 * Each function converts the caller's QLL-layout arrays to the oracle's
 * natural order (orc_from_qll, which follows lib/layout.c:120-205), computes
 * in FP64 and converts back.  Outputs on a subset ("even"/"odd") write only
 * that parity; the other sites keep what the caller passed, as the QDP
 * pack/unpack of qopstag.c leaves them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#include "oracle.h"
#include "qop.h"

typedef QOP_REAL real_t;

static void need_layout(const Layout *l, OrcLayout *ol)
{
  if (!l || !l->physGeom || !l->innerGeom) {
    fprintf(stderr, "qop: the Layout is not set up (call stagDslashSetup first)\n");
    exit(1);
  }
  if (l->nDim != 4 || orc_layout_setup(ol, l->physGeom, l->innerGeom)) {
    fprintf(stderr, "qop: unsupported layout\n");
    exit(1);
  }
}

/* QLL-layout array of nelem reals per site -> natural FP64 [V][nelem] */
static double *to_nat(const Layout *l, const real_t *q, int nelem)
{
  OrcLayout ol;
  need_layout(l, &ol);
  int64_t n = (int64_t)ol.nSites * nelem;
  double *t = malloc(n * sizeof(double)), *nat = malloc(n * sizeof(double));
  for (int64_t i = 0; i < n; i++) t[i] = q[i];
  orc_from_qll(&ol, t, nat, nelem);
  free(t);
  return nat;
}

/* natural FP64 -> QLL-layout array, sites of parity par only (-1: all) */
static void from_nat(const Layout *l, const double *nat, real_t *q, int nelem, int par)
{
  OrcLayout ol;
  need_layout(l, &ol);
  int64_t V = ol.nSites, n = V * nelem;
  double *cur = to_nat(l, q, nelem), *t = malloc(n * sizeof(double));
  for (int64_t s = 0; s < V; s++)
    if (par < 0 || orc_parity(l->physGeom, s) == par) memcpy(&cur[s * nelem], &nat[s * nelem], nelem * sizeof(double));
  orc_to_qll(&ol, cur, t, nelem);
  for (int64_t i = 0; i < n; i++) q[i] = (real_t)t[i];
  free(cur);
  free(t);
}

/* the oracle's solves run 1e-6 below the caller's rsq (not below 1e-26), so
 * that they stand for the exact solution */
static double tight(double rsq) { return fmax(fabs(rsq) * 1e-6, 1e-26); }

static int sub_par(const char *sub) { return sub[0] == 'e' ? 0 : sub[0] == 'o' ? 1 : -1; }
static int sub_idx(const char *sub) { return sub[0] == 'e' ? 1 : sub[0] == 'o' ? 2 : 0; }

/* doublestored staggered links from the forward links u[2mu] (and u3[2mu]):
 * QOP builds its own backward links from the forward ones
 * (qopstag.c:287-301), so the backward arrays the caller holds are not read */
static double *stag_links(const Layout *l, real_t *u[8], real_t *u3[8], int *nterms)
{
  int64_t V = l->physVol, n = V * 18;
  double *fl = malloc(4 * n * sizeof(double)), *ll = u3 ? malloc(4 * n * sizeof(double)) : NULL;
  for (int mu = 0; mu < 4; mu++) {
    double *a = to_nat(l, u[2 * mu], 18);
    memcpy(&fl[mu * n], a, n * sizeof(double));
    free(a);
    if (u3) {
      a = to_nat(l, u3[2 * mu], 18);
      memcpy(&ll[mu * n], a, n * sizeof(double));
      free(a);
    }
  }
  *nterms = u3 ? 16 : 8;
  double *links = malloc((size_t)*nterms * n * sizeof(double));
  orc_stag_doublestore(l->physGeom, fl, ll, links);
  free(fl);
  free(ll);
  return links;
}

/* ------------------------------------------------------------ staggered */
static Layout *g_stag_layout;

void qopStagDslashInit(int *argc, char ***argv)
{
  (void)argc;
  (void)argv;
}

void qopStagDslashFini(void) { g_stag_layout = NULL; }

/* one rank: the rank grid is all ones (qopstag.c:39-76 derives it from QDP) */
void qopStagDslashSetup(Layout *l, int *rg)
{
  g_stag_layout = l;
  if (rg)
    for (int i = 0; i < l->nDim; i++) rg[i] = 1;
}

/* xx(x) = yy(x + len*dir): dir = +-(mu+1), forward fetches from x+mu
 * (qopstag.c:207-222, QDP_N_V_eq_sV applied len times) */
void qopShift(Layout *l, real_t *xx, real_t *yy, int nelem, int dir, int len)
{
  int mu = abs(dir) - 1, hop = (dir > 0 ? 1 : -1) * len;
  int64_t V = l->physVol;
  double *y = to_nat(l, yy, nelem), *x = malloc(V * nelem * sizeof(double));
  for (int64_t s = 0; s < V; s++)
    memcpy(&x[s * nelem], &y[orc_neighbor(l->physGeom, s, mu, hop) * nelem], nelem * sizeof(double));
  from_nat(l, x, xx, nelem, -1);
  free(x);
  free(y);
}

static inline double complex mel(const double *m, int r, int c)
{
  return m[2 * (3 * c + r)] + I * m[2 * (3 * c + r) + 1];
}

/* len > 0: xx(x) = mm(x) yy(x+mu); len < 0: xx(x) = mm(x)^+ yy(x-mu)
 * (qopstag.c:224-244: QDP_M_eq_M_times_sM / QDP_M_eq_Ma_times_sM), mu = dir */
void qopTransportM(Layout *l, real_t *xx, real_t *mm, real_t *yy, int dir, int len)
{
  int64_t V = l->physVol;
  double *m = to_nat(l, mm, 18), *y = to_nat(l, yy, 18), *x = malloc(V * 18 * sizeof(double));
  for (int64_t s = 0; s < V; s++) {
    const double *a = &m[s * 18], *b = &y[orc_neighbor(l->physGeom, s, dir, len > 0 ? 1 : -1) * 18];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        double complex z = 0;
        for (int k = 0; k < 3; k++) z += (len > 0 ? mel(a, r, k) : conj(mel(a, k, r))) * mel(b, k, c);
        x[s * 18 + 2 * (3 * c + r)] = creal(z);
        x[s * 18 + 2 * (3 * c + r) + 1] = cimag(z);
      }
  }
  from_nat(l, x, xx, 18, -1);
  free(x);
  free(m);
  free(y);
}

/* fat7 + Naik smearing (qopstag.c:246-283 via QOP_asqtad_create_L_from_G2) */
void qopSmear(QOP_FIELD *fl[], QOP_FIELD *ll[], Fat7Coeffs *coef, QOP_FIELD *u[], QOP_FIELD *ul[])
{
  Layout *l = fl[0]->l;
  int64_t n = (int64_t)l->physVol * 18;
  double *un = malloc(4 * n * sizeof(double)), *uln = malloc(4 * n * sizeof(double));
  double *fo = malloc(4 * n * sizeof(double)), *lo = malloc(4 * n * sizeof(double));
  for (int mu = 0; mu < 4; mu++) {
    double *a = to_nat(l, u[mu]->f, 18);
    memcpy(&un[mu * n], a, n * sizeof(double));
    free(a);
    a = to_nat(l, (ul ? ul : u)[mu]->f, 18);
    memcpy(&uln[mu * n], a, n * sizeof(double));
    free(a);
  }
  double c[6] = {coef->one_link, coef->three_staple, coef->five_staple,
                 coef->seven_staple, coef->lepage, coef->naik};
  orc_fat7(l->physGeom, c, un, uln, fo, lo);
  for (int mu = 0; mu < 4; mu++) {
    from_nat(l, &fo[mu * n], fl[mu]->f, 18, -1);
    if (ll) from_nat(l, &lo[mu * n], ll[mu]->f, 18, -1);
  }
  free(un);
  free(uln);
  free(fo);
  free(lo);
}

/* x = mass*y + D y on sub (QLL's dslash2 convention, which the reference's
 * testdslash.c:148-155 checks against QOP_asqtad_dslash_qdp) */
void qopDslash(Layout *l, real_t *x, real_t *u[8], real_t *u3[8], real_t mass, real_t *y, char *sub)
{
  int nt;
  double *links = stag_links(l, u, u3, &nt);
  double *yn = to_nat(l, y, 6), *xn = calloc((size_t)l->physVol * 6, sizeof(double));
  orc_dslash(l->physGeom, links, nt, mass, yn, xn, sub_idx(sub));
  from_nat(l, xn, x, 6, sub_par(sub));
  free(links);
  free(yn);
  free(xn);
}

/* even sites: x = m (m^2 - D_eo D_oe)^-1 y, the even part of (m + D)^-1 y
 * (qopstag.c:400-459; compared with solve2 at testdslash.c:418-425) */
void qopSolve(Layout *l, real_t *x, real_t *u[8], real_t *u3[8], real_t mass, real_t *y, double rsq,
              char *sub)
{
  (void)sub;
  int nt;
  double *links = stag_links(l, u, u3, &nt);
  double *yn = to_nat(l, y, 6), *xn = calloc((size_t)l->physVol * 6, sizeof(double));
  orc_cg(l->physGeom, links, nt, mass, yn, xn, tight(rsq), 100000, 0, NULL);
  from_nat(l, xn, x, 6, 0);
  free(links);
  free(yn);
  free(xn);
}

/* x[i] = m_i (m_i^2 - D_eo D_oe)^-1 y on even sites (qopstag.c:327-398) */
void qopSolveMulti(Layout *l, real_t *x[], real_t *u[8], real_t *u3[8], double masses[], real_t *y,
                   int nmasses, double rsq, char *sub)
{
  (void)sub;
  int nt;
  double *links = stag_links(l, u, u3, &nt);
  double *yn = to_nat(l, y, 6);
  for (int i = 0; i < nmasses; i++) {
    double *xn = calloc((size_t)l->physVol * 6, sizeof(double));
    orc_cg(l->physGeom, links, nt, masses[i], yn, xn, tight(rsq), 100000, 0, NULL);
    from_nat(l, xn, x[i], 6, 0);
    free(xn);
  }
  free(links);
  free(yn);
}

/* --------------------------------------------------------------- Wilson */
static Layout *g_wilson_layout;

void qopWilsonDslashInit(int *argc, char ***argv)
{
  (void)argc;
  (void)argv;
}

void qopWilsonDslashFini(void) { g_wilson_layout = NULL; }

void qopWilsonDslashSetup(Layout *l) { g_wilson_layout = l; }

/* 4x4 complex t = (1 + sign*g_dir) (dir < 4) or (1 + sign*g5)/2 (dir = 4),
 * and the rows P that the half spinor keeps (lib/Gamma.lua:46-75) */
static void proj_parts(int dir, int sign, double complex t[4][4], int rows[2])
{
  double g[32];
  orc_gamma(dir, g);
  double f = (dir == 4) ? 0.5 : 1.0;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++)
      t[r][c] = f * ((r == c ? 1.0 : 0.0) + sign * (g[2 * (4 * r + c)] + I * g[2 * (4 * r + c) + 1]));
  rows[0] = (dir == 4 && sign < 0) ? 2 : 0;
  rows[1] = rows[0] + 1;
}

/* half spinor x = P t y, 12 reals per site [2 spins][3 colours][re, im]
 * (sprojMat of lib/Gamma.lua:57-68; qopwilson.c:225-234) */
void qopSproj(Layout *l, real_t *x, real_t *y, int dir, int sign)
{
  double complex t[4][4];
  int rows[2];
  proj_parts(dir, sign, t, rows);
  int64_t V = l->physVol;
  double *yn = to_nat(l, y, 24), *xn = malloc(V * 12 * sizeof(double));
  for (int64_t s = 0; s < V; s++)
    for (int h = 0; h < 2; h++)
      for (int c = 0; c < 3; c++) {
        double complex z = 0;
        for (int k = 0; k < 4; k++) z += t[rows[h]][k] * (yn[s * 24 + 6 * k + 2 * c] + I * yn[s * 24 + 6 * k + 2 * c + 1]);
        xn[s * 12 + 6 * h + 2 * c] = creal(z);
        xn[s * 12 + 6 * h + 2 * c + 1] = cimag(z);
      }
  from_nat(l, xn, x, 12, -1);
  free(yn);
  free(xn);
}

/* full spinor x = f t (P t)^+ y with f = 1/2 (dir < 4) or 1 (dir = 4)
 * (sreconMat of lib/Gamma.lua:69-73; qopwilson.c:236-245) */
void qopSrecon(Layout *l, real_t *x, real_t *y, int dir, int sign)
{
  double complex t[4][4], R[4][2];
  int rows[2];
  proj_parts(dir, sign, t, rows);
  double f = (dir == 4) ? 1.0 : 0.5;
  for (int r = 0; r < 4; r++)
    for (int h = 0; h < 2; h++) {
      double complex z = 0;   /* (t (P t)^+)[r][h] = sum_k t[r][k] conj(t[rows[h]][k]) */
      for (int k = 0; k < 4; k++) z += t[r][k] * conj(t[rows[h]][k]);
      R[r][h] = f * z;
    }
  int64_t V = l->physVol;
  double *yn = to_nat(l, y, 12), *xn = malloc(V * 24 * sizeof(double));
  for (int64_t s = 0; s < V; s++)
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 3; c++) {
        double complex z = 0;
        for (int h = 0; h < 2; h++) z += R[r][h] * (yn[s * 12 + 6 * h + 2 * c] + I * yn[s * 12 + 6 * h + 2 * c + 1]);
        xn[s * 24 + 6 * r + 2 * c] = creal(z);
        xn[s * 24 + 6 * r + 2 * c + 1] = cimag(z);
      }
  from_nat(l, xn, x, 24, -1);
  free(yn);
  free(xn);
}

static double *wilson_links(const Layout *l, real_t *u[8])
{
  int64_t n = (int64_t)l->physVol * 18;
  double *un = malloc(4 * n * sizeof(double)), *links = malloc(8 * n * sizeof(double));
  for (int mu = 0; mu < 4; mu++) {
    double *a = to_nat(l, u[2 * mu], 18);
    memcpy(&un[mu * n], a, n * sizeof(double));
    free(a);
  }
  orc_wilson_doublestore(l->physGeom, un, links);
  free(un);
  return links;
}

static void gamma5_nat(double *v, int64_t V)
{
  for (int64_t s = 0; s < V; s++)
    for (int k = 12; k < 24; k++) v[s * 24 + k] = -v[s * 24 + k];
}

/* QLL's wilsonDslash (lib/wilsonDslash.c:272-285) on sub, built from the
 * forward links u[2mu]; sign = -1 gives g5 D g5 (the dagger hopping term) */
void qopWilsonDslash(Layout *l, real_t *x, real_t *u[8], real_t mass, int sign, real_t *y, char *sub)
{
  int64_t V = l->physVol;
  double *links = wilson_links(l, u);
  double *yn = to_nat(l, y, 24), *xn = calloc(V * 24, sizeof(double));
  if (sign < 0) gamma5_nat(yn, V);
  orc_wilson_dslash(l->physGeom, links, mass, yn, xn, sub_idx(sub));
  if (sign < 0) gamma5_nat(xn, V);
  from_nat(l, xn, x, 24, sub_par(sub));
  free(links);
  free(yn);
  free(xn);
}

/* The system wilsonSolve states: (m^2 - D_eo D_oe) x = y on the even sites
 * with D the Wilson Dslash on a subset (the hopping term -H), i.e.
 * (1 - k^2 H_eo H_oe) z = y with k = 1/m, returned as x = m * (z / m^2).
 * (QOPQDP's call at qopwilson.c:392 passes the mass as its kappa; QOPQDP is
 * absent here, so the oracle solves the system above.) */
static void wilson_solve_one(const Layout *l, const double *links, double mass, const double *yn, double *xn,
                             double rsq)
{
  int64_t V = l->physVol;
  orc_wilson_cgne(l->physGeom, links, 1.0 / mass, yn, xn, tight(rsq), 100000, NULL);
  for (int64_t i = 0; i < V * 24; i++) xn[i] /= mass;
}

void qopWilsonSolve(Layout *l, real_t *x, real_t *u[8], real_t mass, real_t *y, double rsq, char *sub)
{
  (void)sub;
  double *links = wilson_links(l, u);
  double *yn = to_nat(l, y, 24), *xn = calloc((size_t)l->physVol * 24, sizeof(double));
  wilson_solve_one(l, links, mass, yn, xn, rsq);
  from_nat(l, xn, x, 24, 0);
  free(links);
  free(yn);
  free(xn);
}

void qopWilsonSolveMulti(Layout *l, real_t *x[], real_t *u[8], double masses[], real_t *y, int nmasses,
                         double rsq, char *sub)
{
  (void)sub;
  double *links = wilson_links(l, u);
  double *yn = to_nat(l, y, 24);
  for (int i = 0; i < nmasses; i++) {
    double *xn = calloc((size_t)l->physVol * 24, sizeof(double));
    wilson_solve_one(l, links, masses[i], yn, xn, rsq);
    from_nat(l, xn, x[i], 24, 0);
    free(xn);
  }
  free(links);
  free(yn);
}
