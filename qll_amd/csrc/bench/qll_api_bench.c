/*
 * Timing of the QLL C API entry points a QLL program calls with host arrays
 * in QLL layout: dslash2D on "all" and on "even", and solve2D, on an asqtad
 * lattice whose links are made through the same API (random SU(3) with
 * staggered phases, smearFat7D, gaugeDoublestoreD), as testdslash.c does
 * (testdslash.c:255-341).  Also the host <-> device copy rates the transfers
 * are bounded by.  Prints one JSON object; bench.py runs it.
 *
 * usage: qll_api_bench Lx Ly Lz Lt nrep qll_mass
 */
#define _POSIX_C_SOURCE 200112L
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define DOUBLE
#include "common.h"
#include "qll_amd.h"

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double hval(long i, int tag)
{
  unsigned long h = (unsigned long)i * 0x9E3779B97F4A7C15ul + (unsigned long)tag * 0xBF58476D1CE4E5B9ul;
  h ^= h >> 31;
  h *= 0x94D049BB133111EBul;
  h ^= h >> 29;
  return (double)(h >> 11) / 9007199254740992.0 - 0.5;
}

int main(int argc, char *argv[])
{
  int ls[4] = {32, 32, 32, 32}, nrep = 20;
  double mass = 0.02;
  if (argc >= 5)
    for (int i = 0; i < 4; i++) ls[i] = atoi(argv[1 + i]);
  if (argc >= 6) nrep = atoi(argv[5]);
  if (argc >= 7) mass = atof(argv[6]);
  Layout layout;
  memset(&layout, 0, sizeof layout);
  layout.nranks = 1;
  stagDslashSetupD(&layout, 4, ls, NULL);
  const int vol = layout.nSites, NI = layout.nSitesInner;

  /* SU(3) links with staggered phases eta_mu(x) = (-1)^(x_0 + .. + x_{mu-1}) */
  Field U[4], FL[4], LL[4], FLb[4], LLb[4];
  for (int mu = 0; mu < 4; mu++) {
    fieldNew(&U[mu], &layout, 18);
    fieldNew(&FL[mu], &layout, 18);
    fieldNew(&LL[mu], &layout, 18);
    fieldNew(&FLb[mu], &layout, 18);
    fieldNew(&LLb[mu], &layout, 18);
    for (long i = 0; i < 18L * vol; i++) U[mu].f[i] = hval(i, mu + 1);
  }
  Subset all;
  layoutSubset(&all, &layout, "all");
  beginParallelSubset(all) {
    for (int mu = 0; mu < 4; mu++) makeUD(U[mu].f, 3, ti0, ti1);
  }
  endParallelSubset;
  for (int o = 0; o < layout.nSitesOuter; o++)
    for (int ii = 0; ii < NI; ii++) {
      int x[4];
      LayoutIndex li = {0, o * NI + ii};
      layoutCoord(&layout, x, &li);
      for (int mu = 0; mu < 4; mu++) {
        double *m = U[mu].f + 18L * NI * o + ii;
        double complex a[9];
        for (int e = 0; e < 9; e++) a[e] = m[NI * 2 * e] + I * m[NI * (2 * e + 1)];
        /* column-major: a[r + 3c]; det over columns */
        double complex det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                             a[6] * (a[1] * a[5] - a[4] * a[2]);
        double complex ph = cexp(-I * carg(det) / 3.0);
        int eta = 0;
        for (int nu = 0; nu < mu; nu++) eta += x[nu];
        if (eta & 1) ph = -ph;
        for (int e = 0; e < 9; e++) {
          double complex z = a[e] * ph;
          m[NI * 2 * e] = creal(z);
          m[NI * (2 * e + 1)] = cimag(z);
        }
      }
    }
  Fat7Coeffs coef = {5.0 / 8, -1.0 / 16, 1.0 / 64, -1.0 / 384, 0.0, -1.0 / 24};
  Field *up[4] = {&U[0], &U[1], &U[2], &U[3]};
  Field *flp[4] = {&FL[0], &FL[1], &FL[2], &FL[3]}, *llp[4] = {&LL[0], &LL[1], &LL[2], &LL[3]};
  Field *flbp[4] = {&FLb[0], &FLb[1], &FLb[2], &FLb[3]}, *llbp[4] = {&LLb[0], &LLb[1], &LLb[2], &LLb[3]};
  double t0 = now();
  smearFat7(flp, llp, &coef, up, up);
  double t_fat7 = now() - t0;
  gaugeDoublestore(flbp, flp, (int[4]){0, 1, 2, 3}, (int[4]){-1, -1, -1, -1}, 4);
  gaugeDoublestore(llbp, llp, (int[4]){0, 1, 2, 3}, (int[4]){-3, -3, -3, -3}, 4);
  double *u[8], *u3[8];
  for (int mu = 0; mu < 4; mu++) {
    for (long i = 0; i < 18L * vol; i++) {
      FLb[mu].f[i] = -FLb[mu].f[i];
      LLb[mu].f[i] = -LLb[mu].f[i];
    }
    u[2 * mu] = FL[mu].f;
    u[2 * mu + 1] = FLb[mu].f;
    u3[2 * mu] = LL[mu].f;
    u3[2 * mu + 1] = LLb[mu].f;
  }
  StaggeredLinksD sla, sle;
  t0 = now();
  createStaggeredLinksD(&sla, &layout, u, u3, "all");
  double t_links = now() - t0;
  createStaggeredLinksD(&sle, &layout, u, u3, "even");
  double *x = myalloc(sizeof(double) * 6 * vol), *y = myalloc(sizeof(double) * 6 * vol);
  for (long i = 0; i < 6L * vol; i++) {
    y[i] = hval(i, 99);
    x[i] = 0;
  }
  double ms[2];
  StaggeredLinksD *sl[2] = {&sla, &sle};
  for (int k = 0; k < 2; k++) {
    for (int w = 0; w < 3; w++) dslash2D(sl[k], x, mass, y);
    t0 = now();
    for (int r = 0; r < nrep; r++) dslash2D(sl[k], x, mass, y);
    ms[k] = (now() - t0) / nrep * 1e3;
  }
  StaggeredSolverD ss;
  createStaggeredSolverD(&ss, &layout, u, u3, "even");
  ss.noscale = 1;
  t0 = now();
  int its = solve2D(&ss, x, mass, y, 1e-20, 50000, "even");
  double t_solve = now() - t0;
  double rates[4] = {0, 0, 0, 0};
  size_t vbytes = sizeof(double) * 6 * (size_t)vol;
  qamd_host_copy_rate(vbytes, 10, rates);
  printf("{\"lattice\": [%d, %d, %d, %d], \"vlen\": %d, \"fat7_s\": %.4f, \"createStaggeredLinks_s\": %.4f, "
         "\"dslash2D_all_ms\": %.4f, \"dslash2D_all_host_bytes\": %zu, \"dslash2D_even_ms\": %.4f, "
         "\"dslash2D_even_host_bytes\": %zu, \"solve2D_s\": %.4f, \"solve2D_iterations\": %d, "
         "\"copy_gbs\": {\"pinned_h2d\": %.2f, \"pinned_d2h\": %.2f, \"pageable_h2d\": %.2f, "
         "\"pageable_d2h\": %.2f}, \"copy_bytes\": %zu}\n",
         ls[0], ls[1], ls[2], ls[3], NI, t_fat7, t_links, ms[0], 2 * vbytes, ms[1], vbytes, t_solve, its,
         rates[0], rates[1], rates[2], rates[3], vbytes);
  freeStaggeredSolverD(&ss);
  freeStaggeredLinksD(&sla);
  freeStaggeredLinksD(&sle);
  return 0;
}
