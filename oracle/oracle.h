/*
 * FP64 CPU oracle for the QLL Dslash-and-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (qll_amd/) links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker and as the timed CPU baseline.
 *
 * Conventions (natural layout, independent of QLL's SIMD layout):
 *   - lattice extents L[0..3], x[0] fastest: lex = x0 + L0*(x1 + L1*(x2 + L2*x3))
 *   - colour vector: 6 doubles per site (re,im of colours 0..2)
 *   - 3x3 matrix: 18 doubles, column major, element (r,c) at 2*(3c+r)
 *     (lib/vla.c:9-11, PeqMtimesX index jl=2*(j+l*nc) at lib/vla.c:942)
 *   - thin gauge field: u[mu][site][18], mu = 0..3
 *   - staggered link set: links[term][site][18] with QLL's term order
 *     term 2k = u[k] (1-hop), term 2k+1 = u3[k] (3-hop), k = 2mu (from x+mu),
 *     k = 2mu+1 (from x-mu)   (lib/stagDslash.c:438-448, 820-826)
 *   - Wilson spinor: 24 doubles per site, spin-major (lib/vlaw.c:10-12)
 */
#ifndef QLL_ORACLE_H
#define QLL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based random numbers (layout independent) ---- */
uint64_t orc_rng64(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx);
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx);
void orc_gauss2(uint64_t seed, uint64_t stream, uint64_t site, uint64_t pair,
                double *g0, double *g1);

/* ---- lattice helpers ---- */
int64_t orc_volume(const int L[4]);
int64_t orc_neighbor(const int L[4], int64_t s, int mu, int hop);
int orc_parity(const int L[4], int64_t s);

/* ---- gauge field generation ---- */
void orc_su3_from_gauss(double *m /*18*/, const double *g /*18*/);
void orc_random_su3_field(const int L[4], uint64_t seed, double *u /*[4][V][18]*/);
void orc_apply_stag_phases(const int L[4], double *u /*[4][V][18]*/);
void orc_gauss_vector(const int L[4], uint64_t seed, uint64_t stream, int parity,
                      int nreal, double *v /*[V][nreal]*/);

/* ---- fat7 + Naik (lib/gauge.c:264-416) ---- */
void orc_fat7(const int L[4], const double coef[6] /*one,three,five,seven,lepage,naik*/,
              const double *u, const double *ulong, double *fl, double *ll);
void orc_stag_doublestore(const int L[4], const double *fl, const double *ll,
                          double *links /*[16][V][18]*/);
void orc_wilson_doublestore(const int L[4], const double *u, double *links /*[8][V][18]*/);
void orc_plaquette(const int L[4], const double *u, double plaq[6]);

/* ---- staggered operator (lib/stagDslash.c) ---- */
void orc_dslash(const int L[4], const double *links, int nterms, double mass,
                const double *y, double *x, int sub /*0 all,1 even,2 odd*/);
void orc_dslash_fb(const int L[4], const double *fl /*[4][V][18]*/, const double *ll /*or NULL*/,
                   double mass, const double *y, double *x, int sub);
double orc_normal_op(const int L[4], const double *links, int nterms, double mass,
                     const double *p, double *Ap, double *tmp);
int orc_cg(const int L[4], const double *links, int nterms, double mass,
           const double *b, double *x, double rsq, int maxits, int noscale,
           double *final_rsq);
int orc_cgms(const int L[4], const double *links, int nterms, const double *masses,
             int nmasses, const double *b, double **x, double rsq, int maxits,
             int noscale);
double orc_norm2_parity(const int L[4], const double *v, int nreal, int parity);

/* ---- Wilson operator (lib/wilsonDslash.c, lib/Gamma.lua) ---- */
void orc_gamma(int mu, double g[32] /*4x4 complex, row major*/);
void orc_wilson_dslash(const int L[4], const double *links8, double mass,
                       const double *y, double *x, int sub);
void orc_wilson_meo(const int L[4], const double *links8, double kappa,
                    const double *y, double *x, int dagger, double *tmp);
int orc_wilson_cgne(const int L[4], const double *links8, double kappa,
                    const double *b, double *x, double rsq, int maxits,
                    double *final_rsq);

/* ---- QLL host layout (lib/layout.c:21-205) ---- */
typedef struct {
  int nd;
  int physGeom[4], innerGeom[4], outerGeom[4];
  int nSites, nSitesOuter, nSitesInner, nEvenOuter, innerCb, innerCbDir;
} OrcLayout;
int orc_layout_setup(OrcLayout *l, const int L[4], const int inner[4]);
int orc_layout_index(const OrcLayout *l, const int coords[4]);
void orc_layout_coord(const OrcLayout *l, int coords[4], int index);
void orc_to_qll(const OrcLayout *l, const double *nat, double *qll, int nelem);
void orc_from_qll(const OrcLayout *l, const double *qll, double *nat, int nelem);

/* ---- timing helper for the CPU baseline (OpenMP) ---- */
double orc_time_dslash(const int L[4], const double *links, int nterms,
                       const double *y, double *x, int nrep, int *nthreads);

#ifdef __cplusplus
}
#endif
#endif
