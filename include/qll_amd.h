/*
 * qll_amd.h -- device-resident C ABI of libqll_amd (MI355X / gfx950).
 *
 * The QLL drop-in (qll.h, solvers.h) is built on this layer; it is also the
 * interface for callers that keep their fields in HBM between calls
 * (benchmarks, batched solvers, mixed precision).  Plain C
 * types only: opaque handles, pointers and sizes.
 *
 * Which reference interface each group replaces is noted per group; entry
 * points marked "no reference counterpart" are extensions for device-resident
 * callers.
 *
 * Errors: functions returning int give 0 on success and -1 on failure;
 * functions returning a handle give NULL on failure.  qamd_last_error()
 * describes the most recent failure.  Nothing here ever falls back to a host
 * computation: when no usable GPU or kernel image is present every entry point
 * fails.
 */
#ifndef QLL_AMD_H
#define QLL_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage precisions: IEEE double, single, half (FP16), and 16-bit fixed
 * point (Q16: int16 mantissas, vectors with an FP32 max-norm per site, links
 * with an FP32 scale per term; 15-bit resolution at FP16's bandwidth) */
enum { QAMD_F64 = 0, QAMD_F32 = 1, QAMD_F16 = 2, QAMD_Q16 = 3 };
enum { QAMD_EVEN = 0, QAMD_ODD = 1 };

typedef struct qamd_lattice qamd_lattice;   /* 4-d periodic lattice       */
typedef struct qamd_field qamd_field;       /* one parity, nrhs x ncomp    */
typedef struct qamd_gauge qamd_gauge;       /* 4 directions of 3x3 links   */
typedef struct qamd_links qamd_links;       /* doublestored Dslash links   */

/* ---- runtime (replaces the OpenMP team set-up of thread.h:20-29 and the
 *      aligned allocator lib/common.c:9-28 for device data) ---- */
const char *qamd_last_error(void);
int qamd_init(void);
int qamd_sync(void);
/* Device blocks of 8 MB or more that the library frees (work fields of the
 * solvers, fields the caller destroys) are kept, up to 48 GB, and handed out
 * again for allocations of the same size: repeated solves then skip
 * hipMalloc, which has taken up to ~1 s for multi-GB blocks.  This frees
 * them (no reference counterpart; a failed allocation does it too). */
int qamd_release_cache(void);
int qamd_device_name(char *buf, int len);
void *qamd_malloc(size_t bytes);
void qamd_free(void *p);
int qamd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int qamd_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* ---- lattice (replaces layoutSetup lib/layout.c:21-80 for the device
 *      layout; all extents even and >= 4) ---- */
qamd_lattice *qamd_lattice_create(const int dims[4]);
void qamd_lattice_destroy(qamd_lattice *lat);
int64_t qamd_lattice_sites(const qamd_lattice *lat);

/* ---- fields (replaces fieldNew/fieldFree/fieldZero, lib/field.c:3-37).
 *      ncomp = 3 (staggered colour vector) or 12 (Wilson spinor). ---- */
qamd_field *qamd_field_create(qamd_lattice *lat, int prec, int ncomp, int nrhs, int parity);
void qamd_field_destroy(qamd_field *f);
int qamd_field_zero(qamd_field *f);
int qamd_field_copy(qamd_field *dst, const qamd_field *src);  /* converts precision */
/* host <-> device in lexicographic full-lattice order (x fastest), nreal =
 * 2*ncomp doubles per site per rhs, rhs slowest; only the field's parity is
 * read/written */
int qamd_field_from_lex(qamd_field *f, const double *host);
int qamd_field_to_lex(const qamd_field *f, double *host);
/* Gaussian source, unit variance per complex entry, counter-based RNG keyed by
 * (seed, stream+rhs, global lexicographic site) -- no reference counterpart */
int qamd_field_gaussian(qamd_field *f, uint64_t seed, uint64_t stream);
/* squared norm (per rhs) and real dot, accumulated in FP64 (replaces
 * fieldNorm2 lib/field.c:50-63 and r_veq_norm2_X lib/vla.c:656-700) */
int qamd_field_norm2(const qamd_field *f, double *out /* [nrhs] */);
int qamd_field_redot(const qamd_field *a, const qamd_field *b, double *out /* [nrhs] */);
/* y = a*x + b*y (all rhs), precisions may differ; replaces the BLAS-1 set
 * X_veq_X_plus_r_times_X / X_eq_scale_plus_X (lib/vla.c:760-825) */
int qamd_field_axpby(double a, const qamd_field *x, double b, qamd_field *y);

/* ---- QLL host layout <-> device (replaces the layout/shift machinery of
 *      lib/layout.c:120-205; reads innerGeom/outerGeom/innerCb of the QLL
 *      Layout, so any VLEN is accepted) ---- */
typedef struct {
  int physGeom[4], innerGeom[4], outerGeom[4];
  int nSitesOuter, nSitesInner, nEvenOuter, innerCb, innerCbDir;
} qamd_qll_layout;
/* host array in QLL order a[(nelem*NI)*oi + NI*e + ii], element type double
 * (host_prec QAMD_F64) or float (QAMD_F32); nelem = 2*ncomp */
int qamd_field_from_qll(qamd_field *f, const qamd_qll_layout *l, const void *host, int host_prec);
int qamd_field_to_qll(const qamd_field *f, const qamd_qll_layout *l, void *host, int host_prec);
/* host array in a caller's site order: site i of the host array has the
 * coordinates coords[4*i .. 4*i+3] (x fastest convention: coords[4*i+mu] is
 * the mu coordinate) and its nelem = 2*ncomp reals at host[i*nelem ...].
 * The table lists every site of the field's parity (sites of the other
 * parity may be listed too and are skipped; an export leaves them as they
 * are).  This is the site order a QDP/QOP host code exposes: it builds the
 * table once with QDP_get_coords_L for each of its sites, the loop of
 * toQDP_F / fromQDP_F (embedqop.c:57-90, qopstag.c:71-103).  Errors: a
 * coordinate out of range, two sites with the same coordinates, or a site of
 * the parity missing.  nrhs = 1; host_prec QAMD_F64 or QAMD_F32. */
int qamd_field_from_sites(qamd_field *f, const int *coords, int64_t nsites, const void *host, int host_prec);
int qamd_field_to_sites(const qamd_field *f, const int *coords, int64_t nsites, void *host, int host_prec);

/* ---- gauge fields (device), FP64 ---- */
qamd_gauge *qamd_gauge_create(qamd_lattice *lat);
void qamd_gauge_destroy(qamd_gauge *u);
/* host [4][V][18] lexicographic, column-major 3x3 */
int qamd_gauge_from_lex(qamd_gauge *u, const double *host);
int qamd_gauge_to_lex(const qamd_gauge *u, double *host);
/* from four QLL-layout arrays u[mu] (18 reals per site) */
int qamd_gauge_from_qll(qamd_gauge *u, const qamd_qll_layout *l, const void *const host[4], int host_prec);
int qamd_gauge_to_qll(const qamd_gauge *u, const qamd_qll_layout *l, void *const host[4], int host_prec);
/* four arrays host[mu] in a caller's site order (as qamd_field_from_sites;
 * the table lists every site), 18 reals per site; row_major = 1 for matrices
 * stored row by row as QDP does (QLL's are column-major; the reference
 * converts with QDP_M_eq_transpose_M, qopstag.c:123-135) */
int qamd_gauge_from_sites(qamd_gauge *u, const int *coords, int64_t nsites, const void *const host[4], int host_prec,
                          int row_major);
int qamd_gauge_to_sites(const qamd_gauge *u, const int *coords, int64_t nsites, void *const host[4], int host_prec,
                        int row_major);
/* seeded SU(3): complex Gaussian entries -> Gram-Schmidt columns (each
 * projection removed twice, so every matrix is unitary to a few eps) -> det = 1
 * (no reference counterpart; column order of makeU lib/vla.c:1635-1654) */
int qamd_gauge_random_su3(qamd_gauge *u, uint64_t seed);
/* multiply U_mu(x) by eta_mu(x) = (-1)^(sum_{nu<mu} x_nu) */
int qamd_gauge_stag_phases(qamd_gauge *u);
/* fat7 + Naik smearing (replaces smearFat7 lib/gauge.c:264-416);
 * coef = {one_link, three_staple, five_staple, seven_staple, lepage, naik};
 * ll may be NULL */
int qamd_fat7(qamd_gauge *fl, qamd_gauge *ll, const double coef[6],
              const qamd_gauge *u, const qamd_gauge *ulong);
/* plaquettes in the pair order of gaugePlaq (lib/gauge.c:114-183) */
int qamd_gauge_plaquette(const qamd_gauge *u, double plaq[6]);

/* ---- Dslash link sets (replaces createStaggeredLinks lib/stagDslash.c:422-462,
 *      createStaggeredLinksFb lib/stagDslash.c:464-519 and createWilsonLinks
 *      lib/wilsonDslash.c:181-202) ----
 * kind: 0 = asqtad/staggered, doublestored (8 or 16 terms);
 *       1 = Wilson (8 terms);
 *       2 = asqtad/staggered, forward links only (4 fat or 8 = fat + long
 *           terms): half the storage of kind 0; the Dslash forms each
 *           backward term -U_mu(x-mu)^+ y(x-mu) (long: x-3mu) from the other
 *           parity's forward link, as dslashFb2T (lib/stagDslash.c:110-176).
 *           Single-rhs operators and solvers only (no block kernels). */
qamd_links *qamd_links_create(qamd_lattice *lat, int kind, int prec, int nterms);
void qamd_links_destroy(qamd_links *l);
/* host [nterms][V][18] lexicographic, QLL term order (kind 2: U_0..U_3,
 * then L_0..L_3) */
int qamd_links_from_lex(qamd_links *l, const double *host);
/* QLL arrays: staggered u[8] (+ u3[8] for 16 terms, term 2k=u[k], 2k+1=u3[k]);
 * Wilson u[8]; forward-only sets read the forward entries u[2mu] (and
 * u3[2mu]) only */
int qamd_links_from_qll(qamd_links *l, const qamd_qll_layout *lay, const void *const u[],
                        const void *const u3[], int host_prec);
/* doublestore from device fat/long links: t[2mu]=F(x), t[2mu+1]=sgn*F(x-mu)^+
 * (sgn = -1 staggered, +1 Wilson), long links likewise with x-3mu; a
 * forward-only set stores F_mu(x) and L_mu(x) as they are */
int qamd_links_from_gauge(qamd_links *l, const qamd_gauge *fat, const qamd_gauge *lng);
/* copy with precision conversion (same kind and term count) */
int qamd_links_convert(qamd_links *dst, const qamd_links *src);
/* 1 if the Dslash on these links rebuilds third rows instead of loading
 * them: set when every long term of a 16-term (or forward-only 8-term) FP64
 * staggered link set (the
 * case for fat7/Naik links built from SU(3) gauge fields with phases), or
 * every term of an FP64 Wilson set, is a scale times a unitary matrix to
 * 1e-12, after each import, doublestore or conversion.  FP32/FP16/Q16 copies
 * converted from such a set (and 16-bit sets imported or doublestored
 * through FP64) inherit it; FP32 sets imported directly, and links that are
 * not of that form (e.g. hash-filled test fields), are read in full.  The
 * Wilson Dslash rebuilds for one or two rhs only (with more, the rhs threads
 * of a site would each repeat the arithmetic).  No counterpart in QLL; the
 * reconstruction is QUDA's two-row compression, QLL's next-step list names
 * it for the long links. */
int qamd_links_reconstructs(const qamd_links *l);
/* device bytes held by a link set (both parities, with the determinant-sign
 * bytes of a rebuilding set); no counterpart in QLL */
size_t qamd_links_bytes(const qamd_links *l);

/* ---- Dslash (replaces dslash2 lib/stagDslash.c:531-544 and wilsonDslash
 *      lib/wilsonDslash.c:272-285 for one output parity) ----
 * out = a*w + b*D in, on out's parity; w may be NULL (then a is ignored).
 * norm2_out (may be NULL): receives |out|^2 summed over all rhs, FP64
 * accumulated.  Staggered fields of nrhs > 1 (<= 16, FP64 <= 8) run the
 * block kernel, which loads every link once for all rhs (no reference
 * counterpart: QLL runs one dslash2 per source). */
int qamd_dslash(const qamd_links *l, qamd_field *out, const qamd_field *in,
                double a, const qamd_field *w, double b, double *norm2_out);

/* Allocation-free form of qamd_dslash for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc); no counterpart in the reference.  Same arguments
 * and results as qamd_dslash for staggered links and single-rhs fields
 * (every precision).  The |out|^2 reduction uses a fixed device array of
 * 1 MiB per GPU (held until the process ends) instead of library scratch:
 * at most 2^24 sites per parity (lattices up to 2^25 sites, e.g. 64^4);
 * larger lattices, blocks of rhs and Wilson links are rejected with an
 * error.  Needs qamd_init to have run; calls no hipMalloc/hipFree.  Calls
 * that overlap on one GPU (from several host threads) race on the array. */
int qamd_dslash_noalloc(const qamd_links *l, qamd_field *out, const qamd_field *in,
                        double a, const qamd_field *w, double b, double *norm2_out);

/* ---- normal operator (replaces dslashp2 lib/stagDslash.c:546-575):
 *      tmp = D_oe p, Ap = m^2 p - D_eo tmp, pAp = m^2|p|^2 + |tmp|^2 ---- */
int qamd_stag_normal(const qamd_links *l, qamd_field *Ap, const qamd_field *p,
                     qamd_field *tmp, double mass, double *pAp);

/* ---- solvers ---- */
typedef struct {
  int iterations;          /* total sloppy + precise iterations          */
  int reliable_updates;
  double true_rsq;         /* |b - A x|^2 / |b|^2 in FP64 at exit          */
  double iter_rsq;         /* the recurrence's own relative residual^2    */
  double seconds;
  int converged;
  int restarts;            /* search-direction restarts at reliable updates */
  int escalations;         /* 1 if stalled 16-bit sloppy links were replaced by FP32 */
} qamd_solve_info;

/* QLL CG semantics (cgSolveT lib/cg.c:43-122 with dslashp2): solves
 * (m^2 - D_eo D_oe) x = b on even sites to |r|^2 < rsq |b|^2 (iterated),
 * all vectors at the links' precision (FP64 or FP32; 16-bit links are
 * rejected: they run inside qamd_stag_cg_mixed), x starts at 0.  Not
 * rescaled by mass. */
int qamd_stag_cg(const qamd_links *l, qamd_field *x, const qamd_field *b, double mass,
                 double rsq, int maxits, qamd_solve_info *info);
/* Mixed precision CG with reliable updates (no reference counterpart):
 * outer FP64 (links l64, x, b), inner sloppy links ls (FP32, FP16 or Q16,
 * escalating to FP32 when 16-bit links stall; doublestored or forward-only,
 * independently of l64); stops on the FP64 true residual
 * |b - A x|^2 <= rsq |b|^2.  The inner iteration is the single-reduction
 * (Chronopoulos-Gear) CG with FP32 vectors, two launches per iteration,
 * restarted at every reliable update.  delta: reliable-update threshold on
 * |r| relative to its maximum since the last update.  x is overwritten (the
 * solve starts from 0). */
int qamd_stag_cg_mixed(const qamd_links *l64, const qamd_links *ls, qamd_field *x,
                       const qamd_field *b, double mass, double rsq, int maxits,
                       double delta, qamd_solve_info *info);
/* Absolute error of the sloppy normal operator of the block and batched
 * mixed solvers (input and odd temporary in the links' format, FP32
 * output), |(A_s - A) v|/|v| for a Gaussian v against A = m^2 - D_eo D_oe in
 * FP64 (no reference counterpart); the single-rhs mixed CG rounds only the
 * odd temporary, so its operator's error is at most this.  The mixed solvers
 * measure it at set-up (the single-rhs one keeps it with the sloppy links
 * for later solves at the same mass) and use FP32 links in place of 16-bit
 * ones when it is not small against m^2.  ls may be doublestored or
 * forward-only. */
int qamd_stag_sloppy_error(const qamd_links *l64, const qamd_links *ls, double mass, double *eps);
/* Batched mixed-precision CG (no reference counterpart; the batched
 * counterpart of calling solve2, lib/stagDslash.c:594-612, once per source):
 * the nrhs (1..8) sources of the even FP64 colour-vector block b are solved
 * together.  Every inner iteration applies the sloppy normal operator to the
 * whole block in one pair of block-Dslash launches (each link loaded once for
 * all rhs), with per-rhs CG scalars; FP64 reliable updates on the block; each
 * rhs stops on its FP64 true residual |b_r - A x_r|^2 <= rsq |b_r|^2 and
 * leaves the iteration.  info: nrhs entries. */
int qamd_stag_cg_mixed_block(const qamd_links *l64, const qamd_links *ls, qamd_field *x,
                             const qamd_field *b, double mass, double rsq, int maxits,
                             double delta, qamd_solve_info *info);
/* Plain FP64 block CG (no reference counterpart; the baseline the batched
 * mixed solver is measured against): the nrhs (1..8) sources of the even
 * FP64 block b are solved together with textbook CG per rhs (cgSolveT's
 * iteration, lib/cg.c:43-122, unscaled), every vector FP64, the normal
 * operator applied by the block Dslash (FP64 doublestored links loaded once
 * for all rhs); each rhs stops on its iterated |r|^2 < rsq |b_r|^2, checked
 * every 16 iterations.  x starts at 0.  info: nrhs entries. */
int qamd_stag_cg_block(const qamd_links *l, qamd_field *x, const qamd_field *b, double mass, double rsq,
                       int maxits, qamd_solve_info *info);
/* One problem of a heterogeneous batch: its own lattice, gauge field
 * (links), mass and block of 1..8 sources. */
typedef struct {
  const qamd_links *l64;   /* FP64 16-term asqtad links of the problem's lattice */
  const qamd_links *ls;    /* sloppy links (FP32, FP16 or Q16; one precision for all) */
  qamd_field *x;           /* even FP64 colour-vector block (solutions)          */
  const qamd_field *b;     /* even FP64 sources, same nrhs as x                  */
  double mass;             /* solves (mass^2 - D_eo D_oe) x = b                 */
} qamd_stag_problem;
/* Heterogeneous batch of independent mixed-precision CG solves (no reference
 * counterpart; the batched form of calling solve2, lib/stagDslash.c:594-612,
 * once per problem and source): up to 64 problems on lattices of any size
 * advance together, each iteration one fused launch per kernel over all
 * problems (problems grouped by rhs count), FP64 reliable updates per
 * problem, each rhs leaving the iteration when its FP64 true residual
 * |b - A x|^2 <= rsq |b|^2.  use_graph != 0 captures the 16 iterations
 * between host checks in a hipGraph once and replays it.  info: one entry per
 * rhs, problems in order. */
int qamd_stag_cg_mixed_batch(const qamd_stag_problem *probs, int nprob, double rsq, int maxits,
                             double delta, int use_graph, qamd_solve_info *info);
/* QLL multishift CG semantics (cgmsSolveT lib/cgms.c:6-179): shifts[0] is
 * taken as 0, stop on the base residual, vectors at the links' precision. */
int qamd_stag_cgms(const qamd_links *l, qamd_field *const x[], const qamd_field *b,
                   double mass, const double shifts[], int nshift, double rsq,
                   int maxits, qamd_solve_info *info);
/* Multishift with sloppy inner links followed by per-shift mixed-precision
 * refinement to the FP64 true residual rsq on every shift (no reference
 * counterpart). */
int qamd_stag_cgms_mixed(const qamd_links *l64, const qamd_links *ls, qamd_field *const x[],
                         const qamd_field *b, double mass, const double shifts[], int nshift,
                         double rsq, int maxits, qamd_solve_info *info /* [nshift] */);

/* Wilson solver on the even-odd preconditioned operator in the hopping
 * normalisation, Mhat = 1 - kappa^2 H_eo H_oe (H the hopping sum of the Wilson
 * Dslash above), by CG on the normal equations Mhat^+ Mhat x = Mhat^+ b with
 * Mhat^+ = g5 Mhat g5.  All nrhs (<= 16) right-hand sides of b are solved
 * together, sharing link loads.  With sloppy links ls (FP32/FP16) the inner
 * CGNE runs in that precision inside FP64 defect correction; stops when every
 * rhs has |b - Mhat x|^2 <= rsq |b|^2 in FP64.  info: nrhs entries.  Replaces
 * the Wilson solve of lib/wilsonDslash.c:287-364 (see wilsonSolve in qll.h). */
int qamd_wilson_cgne(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b,
                     double kappa, double rsq, int maxits, qamd_solve_info *info);
/* multiply a spinor field by gamma_5 = diag(1,1,-1,-1) (DeGrand-Rossi) */
int qamd_field_gamma5(qamd_field *f);

/* ---- QLL host-layout helpers used by the drop-in API ----
 * Generic shift / transport on host arrays in QLL layout (replaces shiftT
 * lib/stagDslash.c:25-101, the transporter lib/transporter.c:33-193 and Xadj
 * lib/vla.c:827-854): for every site x of the outer-site range
 * [beginOuter, endOuter),
 *     dest(x) = op( [U(x) *] src(x - disp) )  (+ dest(x) if accumulate)
 * nelem reals per site; with u != NULL and td > 0 the source is multiplied by
 * the 3x3 matrix field u at x as r[i,j,k] += u[j,l] v[i,l,k] (td = 2 for
 * complex data, lib/vla.c:1207-1225); adjoint != 0 stores the adjoint of a
 * 3x3 result.  Arrays are element type host_prec (F64 double, F32 float). */
int qamd_qll_shift(const qamd_qll_layout *l, const int disp[4], int nelem, int nc, int td,
                   const void *u, const void *src, void *dest, int host_prec, int beginOuter,
                   int endOuter, int accumulate, int adjoint);
/* sum (squares = 0) or sum of squares (squares = 1) of n host reals, FP64
 * accumulated on the GPU (replaces sumElem_X / norm2_X, lib/vla.c:656-700) */
int qamd_host_sum(const void *host, int64_t n, int host_prec, int squares, double *out);

/* ---- flat device vectors of n reals, prec F64 or F32 (the BLAS-1 of
 *      lib/vla.c:656-825 for solvers driven by host Linop callbacks) ---- */
int qamd_flat_axpby(int64_t n, int prec, double a, const void *x, double b, void *y); /* y = a x + b y */
int qamd_flat_dot(int64_t n, int prec, const void *x, const void *y, double *out);   /* FP64 sum x.y */

/* ---- launch-shape autotuning (no reference counterpart) ----
 * Kernels with several launch shapes (the block Dslash: rhs per launch and
 * XCD block order) time each candidate on first use for a given lattice,
 * precision and block size and keep the fastest.  mode 1: on (the
 * default); 0 or QAMD_TUNE=0 in the environment: off, the default shape is
 * used and the cache cleared; mode >= 2 forces candidate mode - 2 (the last
 * one if fewer) everywhere, for tests.  QAMD_VERBOSE=1 prints each choice. */
int qamd_set_tuning(int mode);

/* ---- link share kept in the MALL (no reference counterpart) ----
 * The staggered Dslash on a 16-term link set with rebuilt long rows loads
 * part of the set (whole 3-hop slot groups, at most the given megabytes per
 * set) through the caches so that it stays in MI355X's 256 MB last-level
 * cache from one call to the next, and streams the rest non-temporally.
 * mb64 applies to FP64 sets (default 52), mb to FP32 and 16-bit sets
 * (default 104); 0 streams every link load, a negative value restores the
 * default (or QAMD_MALL_MB64 / QAMD_MALL_MB from the environment).  A caller
 * whose own vectors need the MALL can set 0.  Results are the same bits for
 * every share. */
int qamd_set_mall_share(int mb64, int mb);

/* ---- timing helpers (hipEvents on the library stream) ---- */
int qamd_time_dslash(const qamd_links *l, qamd_field *out, const qamd_field *in,
                     int nrep, double *ms_per_call);
/* streaming copy of 'bytes' repeated nrep times: read+write GB/s and ms per
 * copy (the measured-bandwidth reference; no reference counterpart) */
int qamd_stream_copy(size_t bytes, int nrep, double *gbs, double *ms_per_copy);
/* host <-> device copy rates of 'bytes' in GB/s (no reference counterpart):
 * rates[0] pinned->device, [1] device->pinned (DMA alone), [2] pageable->
 * device, [3] device->pageable (the library's staging pipeline, the path of
 * the QLL API's host arrays) */
int qamd_host_copy_rate(size_t bytes, int nrep, double rates[4]);

#ifdef __cplusplus
}
#endif
#endif
