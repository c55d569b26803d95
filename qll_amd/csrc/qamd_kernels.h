// Cross-file internal entry points of libqll_amd.
#pragma once
#include "qamd_internal.h"
#include "qamd_objects.h"

namespace qamd {

// runtime.hip: index of the fastest of ncand launch candidates for key,
// measured on first use (run(c, ctx) launches candidate c, 0 on success)
int tune_pick(const char *key, int ncand, int def, int (*run)(int, void *), void *ctx);

// fields.hip: a caller's site table (coordinates of each host site) checked
// and inverted on the host; *inv (device, V ints, lexicographic site -> host
// site index or -1) is freed by the caller with dev_free
int site_table_inverse(const qamd_lattice *lat, const int *coords, int64_t nsites, int need_parity, int **inv);
// fields.hip
const Geom *dev_geom(qamd_lattice *lat);
double *scalar_slot();

// dslash_stag.hip / dslash_wilson.hip
// Reduction: with red_out, {|out|^2, |w|^2} is reduced into partial-sum set
// pset and (finish = true) a finisher stores the totals at red_out; with
// finish = false the partial sums are left in set pset for a caller's finisher;
// rs_fixed: partial sums go there instead of set pset (no scratch reserved).
int stag_dslash_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w,
                       double b, double *red_out, hipStream_t st, const int *skip = nullptr, int pset = 0,
                       bool finish = true, const RedScratch *rs_fixed = nullptr);
// stag_block.hip: staggered Dslash on a block of up to 16 rhs (8 in FP64),
// links loaded once per site; with red, {|out_r|^2, |w_r|^2} for each rhs r
// go to partial-sum set rs (2*nrhs values)
int stag_block_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                      bool red, RedScratch rs, hipStream_t st);
// the same with a coefficient ar[r] per rhs, computing only the rhs whose bit
// is set in act (the others' outputs and sums are left untouched / zero)
int stag_block_launch_m(const qamd_links *l, Field &out, const Field &in, const double *ar, const Field *w,
                        double b, bool red, RedScratch rs, hipStream_t st, unsigned act);
// Scalars of the single-reduction CG step.  Iteration parity ph reads slot
// ph of gam/al/restart and writes slot 1 - ph, so every block of a launch can
// read the previous iteration's values while block 0 stores the new ones;
// block 0 also leaves this iteration's values for the host.
struct CGScal {
  double gam[2], al[2];   // previous |r|^2 and alpha
  int restart[2];         // 1: this iteration takes beta = 0 (p = r, s = A r)
  // keep[ph] 1: the first iteration after a reliable update that keeps p,
  // with kr = r.s_old, kp = p_old.s_old, kq = r.p_old measured (s_old =
  // A p_old); 2: the iteration after it, with rho = r.r_prev and pap =
  // p_prev.A p_prev carried over (see cg_block_scalars)
  int keep[2];
  double kr[2], kp[2], kq[2], rho[2], pap[2];
  double rsq, tt, alpha, beta, maxrr;
  int its;
};
// dslash_stag.hip: fused step of the single-reduction (Chronopoulos-Gear)
// mixed-precision CG, FP32 vectors and sloppy (FP32/FP16/Q16) links.  Every
// block first sums, in fixed order, the partial sums of gamma = |r|^2 (set
// rin, one value per block) and |t|^2 (set tin) and forms alpha and beta
// (identically in every block; no finisher launch), then
// Ar = a r - D t (not stored), p = r + beta p, s = Ar + beta s,
// r -= alpha s, with the |r|^2 partial sums into set rout; x += alpha p is
// applied two iterations at a time, in the odd ones (ph = 1): x += alpha_prev
// p_prev + alpha p, so an even number of steps leaves x complete.
// The sets come from red_set after a red_reserve covering the lattice.
int stag_cgcg_step_launch(const qamd_links *l, const Field &t, Field &r, Field &p, Field &s, Field &x, double a,
                          CGScal *cs, int ph, RedScratch rin, RedScratch tin, RedScratch rout, hipStream_t st,
                          const int *skip);
// stag_block.hip: block CG on (m2[r] - D_eo D_oe) x_r = b_r (one m2 per rhs,
// 1..8 rhs, FP64 x and b), sloppy links ls, FP64 reliable updates on l64;
// x0: start from the x passed in
int stag_cg_block_core(const qamd_links *l64, const qamd_links *ls, qamd_field *x, const qamd_field *b,
                       const double *m2, bool x0, double rsq, int maxits, double delta, qamd_solve_info *info);
// solvers.hip: with 16-bit sloppy links too coarse for this mass (operator
// error not small against m^2), *l32 receives FP32 links made from l64
int sloppy_too_coarse(const qamd_links *l64, const qamd_links *ls, double mass, qamd_links **l32);
// dag: the hermitian conjugate g5 D g5 of the hopping term
int wilson_dslash_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w,
                         double b, double *red_out, hipStream_t st, const int *skip = nullptr, bool dag = false);
// the Wilson solver's fused launches (dslash_wilson.hip): epi 1 stores out =
// a w + b H in with per-rhs |out_r|^2 in slot r of rs; epi 2 applies r_r -=
// alpha_r (a w + b H in)_r (out not stored) with per-rhs |r_r|^2
int wilson_dslash_rhs_launch(const qamd_links *l, Field &out, const Field &in, double a, const Field *w, double b,
                             hipStream_t st, bool dag, int epi, Field *r, const double *alpha, RedScratch rs);

}  // namespace qamd
