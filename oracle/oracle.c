/*
 * FP64 CPU oracle for the QLL Dslash-and-solver path (test infrastructure).
 *
 * Each function restates the reference's algorithm in a plain site loop over
 * a natural (lexicographic, x fastest) full-lattice layout; citations name the
 * reference file:line it follows.  Nothing in qll_amd/ uses this file.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ RNG */
/* SplitMix64 finalizer (Steele, Lea, Flood 2014). */
static inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

uint64_t orc_rng64(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx)
{
  const uint64_t g = 0x9E3779B97F4A7C15ULL;
  uint64_t z = mix64(seed + g * (stream + 1));
  z = mix64(z + g * (site + 1));
  z = mix64(z + g * (idx + 1));
  return z;
}

/* uniform in (0,1] with 53 random bits */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx)
{
  uint64_t z = orc_rng64(seed, stream, site, idx);
  return (double)((z >> 11) + 1) * 0x1.0p-53;
}

/* Box-Muller pair from uniforms 2*pair and 2*pair+1 */
void orc_gauss2(uint64_t seed, uint64_t stream, uint64_t site, uint64_t pair,
                double *g0, double *g1)
{
  double u1 = orc_uniform(seed, stream, site, 2 * pair);
  double u2 = orc_uniform(seed, stream, site, 2 * pair + 1);
  double r = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925 * u2;
  *g0 = r * cos(th);
  *g1 = r * sin(th);
}

/* -------------------------------------------------------------- lattice */
int64_t orc_volume(const int L[4]) { return (int64_t)L[0] * L[1] * L[2] * L[3]; }

static inline void coords_of(const int L[4], int64_t s, int x[4])
{
  for (int i = 0; i < 4; i++) { x[i] = (int)(s % L[i]); s /= L[i]; }
}
static inline int64_t lex_of(const int L[4], const int x[4])
{
  return x[0] + (int64_t)L[0] * (x[1] + (int64_t)L[1] * (x[2] + (int64_t)L[2] * x[3]));
}

int64_t orc_neighbor(const int L[4], int64_t s, int mu, int hop)
{
  int x[4];
  coords_of(L, s, x);
  x[mu] = ((x[mu] + hop) % L[mu] + L[mu]) % L[mu];
  return lex_of(L, x);
}

int orc_parity(const int L[4], int64_t s)
{
  int x[4];
  coords_of(L, s, x);
  return (x[0] + x[1] + x[2] + x[3]) & 1;
}

/* ------------------------------------------------------- 3x3 algebra */
/* column major: element (r,c) at complex index 3c+r */
static inline cplx mget(const double *m, int r, int c)
{
  return m[2 * (3 * c + r)] + I * m[2 * (3 * c + r) + 1];
}
static inline void mset(double *m, int r, int c, cplx z)
{
  m[2 * (3 * c + r)] = creal(z);
  m[2 * (3 * c + r) + 1] = cimag(z);
}
/* d = op(a) * op(b), op = identity or adjoint */
static void mmul(double *d, const double *a, int adja, const double *b, int adjb)
{
  double t[18];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      cplx s = 0;
      for (int k = 0; k < 3; k++) {
        cplx x = adja ? conj(mget(a, k, r)) : mget(a, r, k);
        cplx y = adjb ? conj(mget(b, c, k)) : mget(b, k, c);
        s += x * y;
      }
      mset(t, r, c, s);
    }
  memcpy(d, t, sizeof t);
}
static void madj(double *d, const double *a)
{
  double t[18];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) mset(t, r, c, conj(mget(a, c, r)));
  memcpy(d, t, sizeof t);
}
static cplx mdet(const double *m)
{
  cplx a = mget(m, 0, 0), b = mget(m, 0, 1), c = mget(m, 0, 2);
  cplx d = mget(m, 1, 0), e = mget(m, 1, 1), f = mget(m, 1, 2);
  cplx g = mget(m, 2, 0), h = mget(m, 2, 1), k = mget(m, 2, 2);
  return a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
}

/* Gram-Schmidt of columns in the order of makeUsite (lib/vla.c:1635-1654):
 * normalise column i, then for each j<i remove the projection on column j and
 * renormalise.  Then divide by the principal cube root of det so det = 1.
 * Difference from makeUsite: the projections of column i are removed twice
 * ("twice is enough", Giraud, Langou and Rozloznik 2005).  One pass leaves
 * orthogonality errors of eps * cond(G), up to ~1e-12 among the millions of
 * matrices of a 32^4 field; two passes leave a few eps whatever cond(G). */
void orc_su3_from_gauss(double *m, const double *g)
{
  cplx col[3][3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) col[c][r] = g[2 * (3 * c + r)] + I * g[2 * (3 * c + r) + 1];
  for (int i = 0; i < 3; i++) {
    double n = 0;
    for (int r = 0; r < 3; r++) n += creal(col[i][r] * conj(col[i][r]));
    double ni = 1.0 / sqrt(n);
    for (int r = 0; r < 3; r++) col[i][r] *= ni;
    for (int pass = 0; pass < 2; pass++)
      for (int j = 0; j < i; j++) {
        cplx d = 0;
        for (int r = 0; r < 3; r++) d += conj(col[j][r]) * col[i][r];
        for (int r = 0; r < 3; r++) col[i][r] -= d * col[j][r];
        n = 0;
        for (int r = 0; r < 3; r++) n += creal(col[i][r] * conj(col[i][r]));
        ni = 1.0 / sqrt(n);
        for (int r = 0; r < 3; r++) col[i][r] *= ni;
      }
  }
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) mset(m, r, c, col[c][r]);
  cplx det = mdet(m);
  double ph = atan2(cimag(det), creal(det));
  cplx f = cos(-ph / 3) + I * sin(-ph / 3);
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) mset(m, r, c, mget(m, r, c) * f);
}

/* Link U_mu(x): stream 16+mu, site = global lex index, complex element k
 * (column major) from Gaussian pair k, scaled 1/sqrt(2) per real part. */
void orc_random_su3_field(const int L[4], uint64_t seed, double *u)
{
  int64_t V = orc_volume(L);
#pragma omp parallel for
  for (int64_t s = 0; s < V; s++) {
    for (int mu = 0; mu < 4; mu++) {
      double g[18];
      for (int k = 0; k < 9; k++) {
        double a, b;
        orc_gauss2(seed, 16 + mu, s, k, &a, &b);
        g[2 * k] = a * M_SQRT1_2;
        g[2 * k + 1] = b * M_SQRT1_2;
      }
      orc_su3_from_gauss(&u[(mu * V + s) * 18], g);
    }
  }
}

/* eta_mu(x) = (-1)^(sum_{nu<mu} x_nu), periodic boundaries */
void orc_apply_stag_phases(const int L[4], double *u)
{
  int64_t V = orc_volume(L);
  for (int64_t s = 0; s < V; s++) {
    int x[4];
    coords_of(L, s, x);
    int acc = 0;
    for (int mu = 0; mu < 4; mu++) {
      if (acc & 1)
        for (int i = 0; i < 18; i++) u[(mu * V + s) * 18 + i] = -u[(mu * V + s) * 18 + i];
      acc += x[mu];
    }
  }
}

/* Complex Gaussian vector, unit variance per complex entry, on sites of the
 * given parity (-1 = all); stream chosen by the caller.  Other sites zero. */
void orc_gauss_vector(const int L[4], uint64_t seed, uint64_t stream, int parity,
                      int nreal, double *v)
{
  int64_t V = orc_volume(L);
  for (int64_t s = 0; s < V; s++) {
    int p = orc_parity(L, s);
    for (int k = 0; k < nreal / 2; k++) {
      double a = 0, b = 0;
      if (parity < 0 || p == parity) {
        orc_gauss2(seed, stream, s, k, &a, &b);
        a *= M_SQRT1_2;
        b *= M_SQRT1_2;
      }
      v[s * nreal + 2 * k] = a;
      v[s * nreal + 2 * k + 1] = b;
    }
  }
}

/* ------------------------------------------------------------- fat7 */
typedef struct { const int *L; int64_t V; } Lat;

static void fshift(Lat *g, double *dst, const double *src, int mu, int hop)
{
  /* dst(x) = src(x + hop*mu) */
#pragma omp parallel for
  for (int64_t s = 0; s < g->V; s++)
    memcpy(&dst[s * 18], &src[orc_neighbor(g->L, s, mu, hop) * 18], 18 * sizeof(double));
}

/* Staple sum along nu for a link field X in direction mu, following
 * genStaple (lib/gauge.c:188-262):
 *   upper(x) = U_nu(x) X(x+nu) U_nu(x+mu)^+
 *   lower(x) = U_nu(x-nu)^+ X(x-nu) U_nu(x-nu+mu)
 * out = upper + lower */
static void staple(Lat *g, double *out, const double *u, const double *X, int mu, int nu)
{
  const double *Unu = &u[(int64_t)nu * g->V * 18];
#pragma omp parallel for
  for (int64_t s = 0; s < g->V; s++) {
    double t1[18], t2[18];
    int64_t spn = orc_neighbor(g->L, s, nu, 1);
    int64_t spm = orc_neighbor(g->L, s, mu, 1);
    mmul(t1, &Unu[s * 18], 0, &X[spn * 18], 0);
    mmul(t1, t1, 0, &Unu[spm * 18], 1);
    int64_t smn = orc_neighbor(g->L, s, nu, -1);
    int64_t smnpm = orc_neighbor(g->L, smn, mu, 1);
    mmul(t2, &Unu[smn * 18], 1, &X[smn * 18], 0);
    mmul(t2, t2, 0, &Unu[smnpm * 18], 0);
    for (int i = 0; i < 18; i++) out[s * 18 + i] = t1[i] + t2[i];
  }
}

static void axpy_m(Lat *g, double *y, double a, const double *x)
{
  for (int64_t i = 0; i < g->V * 18; i++) y[i] += a * x[i];
}

/* smearFat7 (lib/gauge.c:264-416): one-link coefficient c1 - 6*lepage
 * (the Lepage trick, :282), 3-, 5-, 7-staples nested, Lepage staple of the
 * 3-staple, and long links naik * U(x)U(x+mu)U(x+2mu) (:369-392). */
void orc_fat7(const int L[4], const double coef[6], const double *u,
              const double *ulong, double *fl, double *ll)
{
  Lat g = {L, orc_volume(L)};
  int64_t n = g.V * 18;
  double c1 = coef[0] - 6 * coef[4], c3 = coef[1], c5 = coef[2], c7 = coef[3];
  double cL = coef[4], naik = coef[5];
  double *S3 = malloc(n * sizeof(double)), *S5 = malloc(n * sizeof(double));
  double *S7 = malloc(n * sizeof(double));
  for (int mu = 0; mu < 4; mu++) {
    double *F = &fl[mu * n];
    const double *Umu = &u[mu * n];
    for (int64_t i = 0; i < n; i++) F[i] = c1 * Umu[i];
    for (int nu = 0; nu < 4; nu++) {
      if (nu == mu) continue;
      staple(&g, S3, u, Umu, mu, nu);
      axpy_m(&g, F, c3, S3);
      if (cL != 0) {
        staple(&g, S5, u, S3, mu, nu);
        axpy_m(&g, F, cL, S5);
      }
      if (c5 != 0 || c7 != 0) {
        for (int rho = 0; rho < 4; rho++) {
          if (rho == mu || rho == nu) continue;
          staple(&g, S5, u, S3, mu, rho);
          axpy_m(&g, F, c5, S5);
          if (c7 != 0) {
            for (int sig = 0; sig < 4; sig++) {
              if (sig == mu || sig == nu || sig == rho) continue;
              staple(&g, S7, u, S5, mu, sig);
              axpy_m(&g, F, c7, S7);
            }
          }
        }
      }
    }
  }
  if (ll) {
    for (int mu = 0; mu < 4; mu++) {
      const double *UL = &ulong[mu * n];
      double *D = &ll[mu * n];
#pragma omp parallel for
      for (int64_t s = 0; s < g.V; s++) {
        int64_t s1 = orc_neighbor(L, s, mu, 1), s2 = orc_neighbor(L, s, mu, 2);
        double t[18];
        mmul(t, &UL[s * 18], 0, &UL[s1 * 18], 0);
        mmul(t, t, 0, &UL[s2 * 18], 0);
        for (int i = 0; i < 18; i++) D[s * 18 + i] = naik * t[i];
      }
    }
  }
  free(S3);
  free(S5);
  free(S7);
}

/* u[2mu] = F_mu(x), u[2mu+1] = -F_mu(x-mu)^+, u3 likewise with x-3mu
 * (testdslash.c:255-285 with the negative adjoint adjM :59-85), stored in the
 * combined 16-term order t[2k]=u[k], t[2k+1]=u3[k]. */
void orc_stag_doublestore(const int L[4], const double *fl, const double *ll, double *links)
{
  int64_t V = orc_volume(L), n = V * 18;
  int nt = ll ? 16 : 8;
  for (int k = 0; k < 8; k++) {
    int mu = k / 2, bwd = k & 1;
    for (int naik = 0; naik < (ll ? 2 : 1); naik++) {
      const double *src = naik ? &ll[mu * n] : &fl[mu * n];
      int term = (nt == 16) ? 2 * k + naik : k;
      double *dst = &links[term * n];
      int hop = naik ? 3 : 1;
#pragma omp parallel for
      for (int64_t s = 0; s < V; s++) {
        if (!bwd) {
          memcpy(&dst[s * 18], &src[s * 18], 18 * sizeof(double));
        } else {
          double t[18];
          madj(t, &src[orc_neighbor(L, s, mu, -hop) * 18]);
          for (int i = 0; i < 18; i++) dst[s * 18 + i] = -t[i];
        }
      }
    }
  }
}

/* Wilson doublestore: u[2mu] = U_mu(x), u[2mu+1] = +U_mu(x-mu)^+
 * (testwilson.c:59-87, gaugeDoublestore lib/gauge.c:34-63) */
void orc_wilson_doublestore(const int L[4], const double *u, double *links)
{
  int64_t V = orc_volume(L), n = V * 18;
  for (int mu = 0; mu < 4; mu++) {
#pragma omp parallel for
    for (int64_t s = 0; s < V; s++) {
      memcpy(&links[(2 * mu) * n + s * 18], &u[mu * n + s * 18], 18 * sizeof(double));
      madj(&links[(2 * mu + 1) * n + s * 18], &u[mu * n + orc_neighbor(L, s, mu, -1) * 18]);
    }
  }
}

/* Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+) / (3 V), for the pairs
 * (nu<mu) in the order of gaugePlaq (lib/gauge.c:114-183). */
void orc_plaquette(const int L[4], const double *u, double plaq[6])
{
  int64_t V = orc_volume(L), n = V * 18;
  int l = 0;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < mu; nu++) {
      double sum = 0;
      for (int64_t s = 0; s < V; s++) {
        double a[18], b[18];
        mmul(a, &u[mu * n + s * 18], 0, &u[nu * n + orc_neighbor(L, s, mu, 1) * 18], 0);
        mmul(b, &u[nu * n + s * 18], 0, &u[mu * n + orc_neighbor(L, s, nu, 1) * 18], 0);
        mmul(a, a, 0, b, 1);
        sum += a[0] + a[8] + a[16];
      }
      plaq[l++] = sum / (3.0 * V);
    }
}

/* ---------------------------------------------------- staggered Dslash */
static inline void mv_acc(double *acc, const double *m, const double *v)
{
  for (int r = 0; r < 3; r++) {
    double re = 0, im = 0;
    for (int c = 0; c < 3; c++) {
      double mr = m[2 * (3 * c + r)], mi = m[2 * (3 * c + r) + 1];
      re += mr * v[2 * c] - mi * v[2 * c + 1];
      im += mr * v[2 * c + 1] + mi * v[2 * c];
    }
    acc[2 * r] += re;
    acc[2 * r + 1] += im;
  }
}

/* dslash2 (lib/stagDslash.c:259-420, 531-544):
 *   x(s) = massFac*mass*y(s) + sum_i t_i(s) y(s - disp_i)
 * massFac = 0 for the even/odd subsets (:434-436).  Term i: k = i/2 (with
 * Naik) or i (without), mu = k/2, from x+mu for even k, x-mu for odd k,
 * hop 3 for the odd terms of the 16-term set. */
void orc_dslash(const int L[4], const double *links, int nterms, double mass,
                const double *y, double *x, int sub)
{
  int64_t V = orc_volume(L), n = V * 18;
  double mf = (sub == 0) ? mass : 0.0;
#pragma omp parallel for
  for (int64_t s = 0; s < V; s++) {
    int p = orc_parity(L, s);
    if (sub == 1 && p != 0) continue;
    if (sub == 2 && p != 1) continue;
    double acc[6];
    for (int i = 0; i < 6; i++) acc[i] = mf * y[s * 6 + i];
    for (int t = 0; t < nterms; t++) {
      int k = (nterms == 16) ? t / 2 : t;
      int hop = (nterms == 16 && (t & 1)) ? 3 : 1;
      int mu = k / 2;
      int dir = (k & 1) ? -hop : hop;
      mv_acc(acc, &links[t * n + s * 18], &y[orc_neighbor(L, s, mu, dir) * 6]);
    }
    for (int i = 0; i < 6; i++) x[s * 6 + i] = acc[i];
  }
}

/* dslashFb2T (lib/stagDslash.c:110-176) on the links of createStaggeredLinksFb
 * (lib/stagDslash.c:464-519): forward links only, fl[mu] (and ll[mu]);
 *   x(s) = massFac*mass*y(s) + sum_mu fl_mu(s) y(s+mu) - fl_mu(s-mu)^+ y(s-mu)
 *                            + sum_mu ll_mu(s) y(s+3mu) - ll_mu(s-3mu)^+ y(s-3mu)
 * (the backward terms by VmeqMAXtimesVXp, :164-170) */
void orc_dslash_fb(const int L[4], const double *fl, const double *ll, double mass,
                   const double *y, double *x, int sub)
{
  int64_t V = orc_volume(L), n = V * 18;
  double mf = (sub == 0) ? mass : 0.0;
#pragma omp parallel for
  for (int64_t s = 0; s < V; s++) {
    int p = orc_parity(L, s);
    if (sub == 1 && p != 0) continue;
    if (sub == 2 && p != 1) continue;
    double acc[6];
    for (int i = 0; i < 6; i++) acc[i] = mf * y[s * 6 + i];
    for (int naik = 0; naik < (ll ? 2 : 1); naik++) {
      const double *u = naik ? ll : fl;
      int hop = naik ? 3 : 1;
      for (int mu = 0; mu < 4; mu++) {
        mv_acc(acc, &u[mu * n + s * 18], &y[orc_neighbor(L, s, mu, hop) * 6]);
        int64_t b = orc_neighbor(L, s, mu, -hop);
        double t[18], a[6] = {0, 0, 0, 0, 0, 0};
        madj(t, &u[mu * n + b * 18]);
        mv_acc(a, t, &y[b * 6]);
        for (int i = 0; i < 6; i++) acc[i] -= a[i];
      }
    }
    for (int i = 0; i < 6; i++) x[s * 6 + i] = acc[i];
  }
}

double orc_norm2_parity(const int L[4], const double *v, int nreal, int parity)
{
  int64_t V = orc_volume(L);
  double s = 0;
  for (int64_t i = 0; i < V; i++) {
    if (parity >= 0 && orc_parity(L, i) != parity) continue;
    for (int k = 0; k < nreal; k++) s += v[i * nreal + k] * v[i * nreal + k];
  }
  return s;
}

/* dslashp2 (lib/stagDslash.c:546-575): tmp_o = D_oe p, Ap_e = m^2 p - D_eo tmp,
 * returns Anorm2 = m^2 |p_e|^2 + |tmp_o|^2.  Odd sites of Ap are untouched. */
double orc_normal_op(const int L[4], const double *links, int nterms, double mass,
                     const double *p, double *Ap, double *tmp)
{
  int64_t V = orc_volume(L);
  orc_dslash(L, links, nterms, 0, p, tmp, 2);
  orc_dslash(L, links, nterms, 0, tmp, Ap, 1);
  double m2 = mass * mass;
  for (int64_t s = 0; s < V; s++) {
    if (orc_parity(L, s)) continue;
    for (int k = 0; k < 6; k++) Ap[s * 6 + k] = m2 * p[s * 6 + k] - Ap[s * 6 + k];
  }
  return m2 * orc_norm2_parity(L, p, 6, 0) + orc_norm2_parity(L, tmp, 6, 1);
}

/* cgSolveT (lib/cg.c:43-122) with solve2's scaling (lib/stagDslash.c:600-623):
 * x=0, r=b, p=0; loop { p += r/rsq; alpha = 1/Anorm2; x += alpha p;
 * r -= alpha Ap; rsq = |r|^2 } while rsq >= rsqin*|b|^2 && its < maxits.
 * Then x <- mass*x unless noscale.  Only even sites are used. */
int orc_cg(const int L[4], const double *links, int nterms, double mass,
           const double *b, double *x, double rsqin, int maxits, int noscale,
           double *final_rsq)
{
  int64_t V = orc_volume(L), n6 = V * 6;
  double *r = calloc(n6, sizeof(double)), *p = calloc(n6, sizeof(double));
  double *Ap = calloc(n6, sizeof(double)), *t = calloc(n6, sizeof(double));
  char *ev = malloc(V);
  for (int64_t s = 0; s < V; s++) ev[s] = (orc_parity(L, s) == 0);
  for (int64_t s = 0; s < V; s++)
    if (ev[s])
      for (int k = 0; k < 6; k++) { x[s * 6 + k] = 0; r[s * 6 + k] = b[s * 6 + k]; }
  double rsq = orc_norm2_parity(L, r, 6, 0), bsq = rsq, rsqstop = rsqin * bsq;
  int itn = 0;
  do {
    itn++;
    double beta = 1 / rsq;
    for (int64_t i = 0; i < n6; i++)
      if (ev[i / 6]) p[i] += beta * r[i];
    double pAp = orc_normal_op(L, links, nterms, mass, p, Ap, t);
    double alpha = 1 / pAp;
    for (int64_t i = 0; i < n6; i++)
      if (ev[i / 6]) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
    rsq = orc_norm2_parity(L, r, 6, 0);
  } while (rsq >= rsqstop && itn < maxits);
  if (final_rsq) *final_rsq = rsq / bsq;
  if (!noscale)
    for (int64_t i = 0; i < n6; i++)
      if (ev[i / 6]) x[i] *= mass;
  free(r); free(p); free(Ap); free(t); free(ev);
  return itn;
}

/* cgmsSolveT (lib/cgms.c:6-179) via solveMulti2 (lib/stagDslash.c:625-657):
 * shifts m2[i] = masses[i]^2 - masses[0]^2 on top of A = masses[0]^2 - Deo Doe,
 * stop on the base residual, x[i] <- masses[i]*x[i] unless noscale. */
int orc_cgms(const int L[4], const double *links, int nterms, const double *masses,
             int ns, const double *in, double **out, double rsqin, int itnlim,
             int noscale)
{
  int64_t V = orc_volume(L), n6 = V * 6;
  double a[ns], b[ns], s[ns], bo[ns], z[ns], zo[ns], zn[ns], shifts[ns];
  for (int i = 0; i < ns; i++) shifts[i] = masses[i] * masses[i] - masses[0] * masses[0];
  char *ev = malloc(V);
  for (int64_t q = 0; q < V; q++) ev[q] = (orc_parity(L, q) == 0);
  double insq = orc_norm2_parity(L, in, 6, 0);
  int iteration = 0;
  if (insq == 0) {
    for (int i = 0; i < ns; i++)
      for (int64_t k = 0; k < n6; k++)
        if (ev[k / 6]) out[i][k] = 0;
    free(ev);
    return 0;
  }
  double *r = calloc(n6, sizeof(double)), *Ap = calloc(n6, sizeof(double));
  double *t = calloc(n6, sizeof(double));
  double *ps[ns];
  for (int i = 0; i < ns; i++) ps[i] = calloc(n6, sizeof(double));
  for (int64_t k = 0; k < n6; k++) r[k] = ev[k / 6] ? in[k] : 0;
  for (int i = 0; i < ns; i++) {
    for (int64_t k = 0; k < n6; k++)
      if (ev[k / 6]) { out[i][k] = 0; ps[i][k] = r[k]; }
    zo[i] = z[i] = 1;
    bo[i] = -1;
    a[i] = 0;
    s[i] = 1;
  }
  double rsqstop = rsqin * insq, rsq = insq, oldrsq;
  while (1) {
    oldrsq = rsq;
    double pkp = orc_normal_op(L, links, nterms, masses[0], ps[0], Ap, t);
    iteration++;
    pkp *= s[0] * s[0];
    if (pkp <= 0) break;
    b[0] = rsq / pkp;
    zn[0] = 1;
    for (int i = 1; i < ns; i++) {
      zn[i] = z[i] * zo[i] * bo[0];
      double c1 = b[0] * a[0] * (zo[i] - z[i]);
      c1 += zo[i] * bo[0] * (1 + shifts[i] * b[0]);
      if (c1 != 0.0) zn[i] /= c1; else zn[i] = 0;
      if (z[i] != 0.0) b[i] = b[0] * zn[i] / z[i]; else zn[i] = b[i] = 0;
    }
    for (int i = 0; i < ns; i++) {
      double tt = b[i] * s[i];
      for (int64_t k = 0; k < n6; k++)
        if (ev[k / 6]) out[i][k] += tt * ps[i][k];
    }
    double t0 = b[0] * s[0];
    for (int64_t k = 0; k < n6; k++)
      if (ev[k / 6]) r[k] -= t0 * Ap[k];
    rsq = orc_norm2_parity(L, r, 6, 0);
    if ((iteration >= itnlim) || (rsqstop <= 0 || rsq < rsqstop)) break;
    a[0] = rsq / oldrsq;
    for (int i = 1; i < ns; i++) {
      double c2 = z[i] * b[0];
      if (c2 != 0.0) a[i] = a[0] * zn[i] * b[i] / c2; else a[i] = 0;
    }
    for (int64_t k = 0; k < n6; k++)
      if (ev[k / 6]) ps[0][k] = a[0] * ps[0][k] + r[k];
    for (int i = 1; i < ns; i++)
      if (s[i] != 0)
        for (int64_t k = 0; k < n6; k++)
          if (ev[k / 6]) ps[i][k] = a[i] * ps[i][k] + zn[i] * r[k];
    for (int i = 0; i < ns; i++) { bo[i] = b[i]; zo[i] = z[i]; z[i] = zn[i]; }
  }
  if (!noscale)
    for (int i = 0; i < ns; i++)
      for (int64_t k = 0; k < n6; k++)
        if (ev[k / 6]) out[i][k] *= masses[i];
  for (int i = 0; i < ns; i++) free(ps[i]);
  free(r); free(Ap); free(t); free(ev);
  return iteration;
}

/* --------------------------------------------------------------- Wilson */
/* DeGrand-Rossi gamma matrices, row-major 4x4 complex (lib/Gamma.lua:7-11):
 * mu=0 -> gamma[1], 1 -> gamma[2], 2 -> gamma[4], 3 -> gamma[8], 4 -> g5 */
void orc_gamma(int mu, double g[32])
{
  static const int re[4][16] = {
    {0,0,0,0, 0,0,0,0, 0,0,0,0, 0,0,0,0},
    {0,0,0,-1, 0,0,1,0, 0,1,0,0, -1,0,0,0},
    {0,0,0,0, 0,0,0,0, 0,0,0,0, 0,0,0,0},
    {0,0,1,0, 0,0,0,1, 1,0,0,0, 0,1,0,0}};
  static const int im[4][16] = {
    {0,0,0,1, 0,0,1,0, 0,-1,0,0, -1,0,0,0},
    {0,0,0,0, 0,0,0,0, 0,0,0,0, 0,0,0,0},
    {0,0,1,0, 0,0,0,-1, -1,0,0,0, 0,1,0,0},
    {0,0,0,0, 0,0,0,0, 0,0,0,0, 0,0,0,0}};
  if (mu < 4) {
    for (int i = 0; i < 16; i++) { g[2 * i] = re[mu][i]; g[2 * i + 1] = im[mu][i]; }
    return;
  }
  /* gamma5 = gamma[15] = g1 g2 g4 g8 (Gamma.lua:12-26 product order) */
  cplx m[16], t[16];
  for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1 : 0;
  for (int d = 0; d < 4; d++) {
    double gd[32];
    orc_gamma(d, gd);
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) {
        cplx s = 0;
        for (int k = 0; k < 4; k++) s += m[4 * r + k] * (gd[2 * (4 * k + c)] + I * gd[2 * (4 * k + c) + 1]);
        t[4 * r + c] = s;
      }
    memcpy(m, t, sizeof m);
  }
  for (int i = 0; i < 16; i++) { g[2 * i] = creal(m[i]); g[2 * i + 1] = cimag(m[i]); }
}

/* wilsonDslash (lib/wilsonDslash.c:21-179, 272-285):
 *   x = massFac*(4+mass)*y - sum_mu [ (1-g_mu) U_mu(x) y(x+mu)
 *                                   + (1+g_mu) U_mu(x-mu)^+ y(x-mu) ]
 * term 2mu uses spinproj{mu}m (sign -1), term 2mu+1 spinproj{mu}p (+1),
 * both subtracted (lib/vlaw0.c.lua:545-649, Gamma.lua:53-75). */
void orc_wilson_dslash(const int L[4], const double *links, double mass,
                       const double *y, double *x, int sub)
{
  int64_t V = orc_volume(L), n = V * 18;
  double mf = (sub == 0) ? (4 + mass) : 0.0;
  double gm[4][32];
  for (int mu = 0; mu < 4; mu++) orc_gamma(mu, gm[mu]);
#pragma omp parallel for
  for (int64_t s = 0; s < V; s++) {
    int p = orc_parity(L, s);
    if (sub == 1 && p != 0) continue;
    if (sub == 2 && p != 1) continue;
    cplx acc[4][3];
    for (int sp = 0; sp < 4; sp++)
      for (int c = 0; c < 3; c++)
        acc[sp][c] = mf * (y[s * 24 + 6 * sp + 2 * c] + I * y[s * 24 + 6 * sp + 2 * c + 1]);
    for (int t = 0; t < 8; t++) {
      int mu = t / 2;
      double sign = (t & 1) ? 1.0 : -1.0;
      int64_t nb = orc_neighbor(L, s, mu, (t & 1) ? -1 : 1);
      const double *m = &links[t * n + s * 18];
      cplx uy[4][3];
      for (int sp = 0; sp < 4; sp++)
        for (int r = 0; r < 3; r++) {
          cplx z = 0;
          for (int c = 0; c < 3; c++)
            z += mget(m, r, c) * (y[nb * 24 + 6 * sp + 2 * c] + I * y[nb * 24 + 6 * sp + 2 * c + 1]);
          uy[sp][r] = z;
        }
      for (int a = 0; a < 4; a++)
        for (int r = 0; r < 3; r++) {
          cplx z = uy[a][r];
          for (int b = 0; b < 4; b++) {
            cplx g = gm[mu][2 * (4 * a + b)] + I * gm[mu][2 * (4 * a + b) + 1];
            z += sign * g * uy[b][r];
          }
          acc[a][r] -= z;
        }
    }
    for (int sp = 0; sp < 4; sp++)
      for (int c = 0; c < 3; c++) {
        x[s * 24 + 6 * sp + 2 * c] = creal(acc[sp][c]);
        x[s * 24 + 6 * sp + 2 * c + 1] = cimag(acc[sp][c]);
      }
  }
}

/* Even-odd preconditioned Wilson operator in the hopping normalisation:
 *   Mhat y_e = y_e - kappa^2 H_eo H_oe y_e,  H = the hopping sum of the
 * QLL Dslash (so D_oe y = -H_oe y with massFac=0).  dagger=1 applies
 * Mhat^+ = g5 Mhat g5.  Only even sites of x are written. */
void orc_wilson_meo(const int L[4], const double *links, double kappa,
                    const double *y, double *x, int dagger, double *tmp)
{
  int64_t V = orc_volume(L);
  double g5[32];
  orc_gamma(4, g5);
  const double *src = y;
  double *yy = NULL;
  if (dagger) {
    /* DeGrand-Rossi g5 is diagonal: diag(1,1,-1,-1) times its phase */
    yy = malloc(V * 24 * sizeof(double));
    for (int64_t s = 0; s < V; s++)
      for (int sp = 0; sp < 4; sp++) {
        double f = g5[2 * (5 * sp)];
        for (int k = 0; k < 6; k++) yy[s * 24 + 6 * sp + k] = f * y[s * 24 + 6 * sp + k];
      }
    src = yy;
  }
  orc_wilson_dslash(L, links, 0, src, tmp, 2);
  orc_wilson_dslash(L, links, 0, tmp, x, 1);
  double k2 = kappa * kappa;
  for (int64_t s = 0; s < V; s++) {
    if (orc_parity(L, s)) continue;
    for (int sp = 0; sp < 4; sp++) {
      double f = dagger ? g5[2 * (5 * sp)] : 1.0;
      for (int k = 0; k < 6; k++) {
        int i = s * 24 + 6 * sp + k;
        x[i] = f * (src[i] - k2 * x[i]);
      }
    }
  }
  free(yy);
}

/* CGNE on Mhat^+ Mhat x = Mhat^+ b (textbook CG), stopping on the true
 * residual |b - Mhat x|^2 <= rsq |b|^2, checked each time the iterated
 * normal-equation residual says it should be. */
int orc_wilson_cgne(const int L[4], const double *links, double kappa,
                    const double *b, double *x, double rsq, int maxits,
                    double *final_rsq)
{
  int64_t V = orc_volume(L), n = V * 24;
  double *r = calloc(n, 8), *p = calloc(n, 8), *q = calloc(n, 8), *t = calloc(n, 8);
  double *tmp = calloc(n, 8), *Ap = calloc(n, 8);
  for (int64_t i = 0; i < n; i++) x[i] = 0;
  orc_wilson_meo(L, links, kappa, b, r, 1, tmp); /* r = M^+ b */
  for (int64_t s = 0; s < V; s++)
    if (orc_parity(L, s))
      for (int k = 0; k < 24; k++) r[s * 24 + k] = 0;
  memcpy(p, r, n * 8);
  double bsq = orc_norm2_parity(L, b, 24, 0);
  double rr = orc_norm2_parity(L, r, 24, 0), rr0 = rr;
  int it = 0;
  double true_rsq = 1;
  while (it < maxits) {
    orc_wilson_meo(L, links, kappa, p, q, 0, tmp);
    double qq = orc_norm2_parity(L, q, 24, 0);
    double alpha = rr / qq;
    orc_wilson_meo(L, links, kappa, q, Ap, 1, tmp);
    for (int64_t s = 0; s < V; s++) {
      if (orc_parity(L, s)) continue;
      for (int k = 0; k < 24; k++) {
        int64_t i = s * 24 + k;
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
      }
    }
    double rrn = orc_norm2_parity(L, r, 24, 0);
    it++;
    if (rrn < rsq * rsq * rr0 || rrn < 1e-30 * rr0) {
      orc_wilson_meo(L, links, kappa, x, t, 0, tmp);
      double e = 0;
      for (int64_t s = 0; s < V; s++) {
        if (orc_parity(L, s)) continue;
        for (int k = 0; k < 24; k++) { double d = b[s * 24 + k] - t[s * 24 + k]; e += d * d; }
      }
      true_rsq = e / bsq;
      if (true_rsq <= rsq) break;
    }
    double beta = rrn / rr;
    rr = rrn;
    for (int64_t s = 0; s < V; s++) {
      if (orc_parity(L, s)) continue;
      for (int k = 0; k < 24; k++) p[s * 24 + k] = r[s * 24 + k] + beta * p[s * 24 + k];
    }
  }
  if (final_rsq) *final_rsq = true_rsq;
  free(r); free(p); free(q); free(t); free(tmp); free(Ap);
  return it;
}

/* ------------------------------------------------------- QLL layout */
/* partition-free restatement of layoutSetup (lib/layout.c:21-80) for one
 * rank; returns 0 on success, -1 where QLL exits ("not enough 2's"),
 * -2 for QLL's "error in cb choice". */
int orc_layout_setup(OrcLayout *l, const int L[4], const int inner[4])
{
  int nd = 4, lvol = 1, ovol = 1, icb = 0, icbd = -1;
  l->nd = nd;
  for (int i = 0; i < nd; i++) {
    l->physGeom[i] = L[i];
    l->innerGeom[i] = inner[i];
    l->outerGeom[i] = L[i] / inner[i];
    lvol *= L[i];
    ovol *= l->outerGeom[i];
    if (inner[i] > 1 && (l->outerGeom[i] & 1) == 1) icb++;
    if (inner[i] == 1 && (l->outerGeom[i] & 1) == 0) icbd = i;
  }
  if (icb == 0) {
    icbd = 0;
  } else {
    if (icbd < 0) return -1;
    icb = l->outerGeom[icbd] / 2;
    if ((icb & 1) == 0) return -2;
  }
  l->nSites = lvol;
  l->nSitesOuter = ovol;
  l->nEvenOuter = ovol - ovol / 2;
  l->nSitesInner = lvol / ovol;
  l->innerCb = icb;
  l->innerCbDir = icbd;
  return 0;
}

/* layoutIndex (lib/layout.c:120-142), single rank */
int orc_layout_index(const OrcLayout *l, const int coords[4])
{
  int nd = l->nd, x[4];
  for (int i = 0; i < nd; i++) x[i] = coords[i];
  int ii = 0;
  for (int i = nd - 1; i >= 0; i--) ii = ii * l->innerGeom[i] + ((x[i] / l->outerGeom[i]) % l->innerGeom[i]);
  int ib = 0;
  for (int i = 0; i < nd; i++) ib += ((x[i] / l->outerGeom[i]) % l->innerGeom[i]) * l->outerGeom[i];
  ib &= 1;
  x[l->innerCbDir] += l->innerCb * ib;
  int oi = 0;
  for (int i = nd - 1; i >= 0; i--) oi = oi * l->outerGeom[i] + (x[i] % l->outerGeom[i]);
  int p = 0;
  for (int i = 0; i < nd; i++) p += coords[i];
  int oi2 = (p & 1) ? (oi + l->nSitesOuter) / 2 : oi / 2;
  return oi2 * l->nSitesInner + ii;
}

/* inverse of orc_layout_index by search-free reconstruction
 * (layoutCoord, lib/layout.c:144-205) */
void orc_layout_coord(const OrcLayout *l, int coords[4], int index)
{
  int nd = l->nd, p = 0, cr[4];
  int ll = index % l->nSitesInner, ib = 0;
  for (int i = 0; i < nd; i++) {
    int w = l->innerGeom[i], wl = l->outerGeom[i], k = ll % w;
    cr[i] = k * wl;
    p += cr[i];
    ll /= w;
    ib += k * wl;
  }
  ib &= 1;
  int ii = index / l->nSitesInner;
  if (ii >= l->nEvenOuter) { ii -= l->nEvenOuter; p++; }
  ii *= 2;
  for (int i = 0; i < nd; i++) {
    int wl = l->outerGeom[i], k = ii % wl;
    if (i == l->innerCbDir) k = (k + l->innerCb * ib) % wl;
    coords[i] = k;
    p += k;
    ii /= wl;
  }
  if (p & 1) {
    for (int i = 0; i < nd; i++) {
      int wl = l->outerGeom[i];
      if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
      coords[i]++;
      if (coords[i] >= wl) {
        coords[i] = 0;
        if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
      } else {
        if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
        break;
      }
    }
  }
  for (int i = 0; i < nd; i++) coords[i] += cr[i];
}

/* QLL host array a[(nelem*NI)*oi + NI*e + ii], index = oi*NI + ii
 * (qopstag.c:90-95) */
void orc_to_qll(const OrcLayout *l, const double *nat, double *qll, int nelem)
{
  int NI = l->nSitesInner;
  for (int64_t s = 0; s < l->nSites; s++) {
    int x[4];
    coords_of(l->physGeom, s, x);
    int idx = orc_layout_index(l, x);
    int oi = idx / NI, ii = idx % NI;
    for (int e = 0; e < nelem; e++) qll[(int64_t)nelem * NI * oi + NI * e + ii] = nat[s * nelem + e];
  }
}

void orc_from_qll(const OrcLayout *l, const double *qll, double *nat, int nelem)
{
  int NI = l->nSitesInner;
  for (int64_t s = 0; s < l->nSites; s++) {
    int x[4];
    coords_of(l->physGeom, s, x);
    int idx = orc_layout_index(l, x);
    int oi = idx / NI, ii = idx % NI;
    for (int e = 0; e < nelem; e++) nat[s * nelem + e] = qll[(int64_t)nelem * NI * oi + NI * e + ii];
  }
}

/* ------------------------------------------------------ CPU baseline */
/* The dslash2T loop nest of lib/stagDslash.c:316-328 (256-site blocks of
 * even output sites, 16 terms each), OpenMP over blocks, one parity. */
double orc_time_dslash(const int L[4], const double *links, int nterms,
                       const double *y, double *x, int nrep, int *nthreads)
{
  int64_t V = orc_volume(L), n = V * 18;
  int64_t nev = 0;
  int64_t *sites = malloc(V * sizeof(int64_t));
  for (int64_t s = 0; s < V; s++)
    if (orc_parity(L, s) == 0) sites[nev++] = s;
  /* neighbour tables, built once like QLL's shift setup */
  int64_t *nb = malloc(nev * nterms * sizeof(int64_t));
  for (int64_t i = 0; i < nev; i++)
    for (int t = 0; t < nterms; t++) {
      int k = (nterms == 16) ? t / 2 : t;
      int hop = (nterms == 16 && (t & 1)) ? 3 : 1;
      nb[i * nterms + t] = orc_neighbor(L, sites[i], k / 2, (k & 1) ? -hop : hop);
    }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
  double t0 = omp_get_wtime();
#else
  double t0 = 0;
#endif
  for (int rep = 0; rep < nrep; rep++) {
#pragma omp parallel for schedule(static)
    for (int64_t b0 = 0; b0 < nev; b0 += 256) {
      int64_t b1 = b0 + 256 < nev ? b0 + 256 : nev;
      for (int64_t i = b0; i < b1; i++) {
        int64_t s = sites[i];
        for (int k = 0; k < 6; k++) x[s * 6 + k] = 0;
      }
      for (int t = 0; t < nterms; t++)
        for (int64_t i = b0; i < b1; i++) {
          int64_t s = sites[i];
          mv_acc(&x[s * 6], &links[t * n + s * 18], &y[nb[i * nterms + t] * 6]);
        }
    }
  }
#ifdef _OPENMP
  double t1 = omp_get_wtime();
#else
  double t1 = 0;
#endif
  free(sites);
  free(nb);
  if (nthreads) *nthreads = nth;
  return (t1 - t0) / nrep;
}
