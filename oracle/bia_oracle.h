/*
 * bia_oracle.h -- CPU restatement of the reference's pairwise BIA hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  It is built into
 * oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  The product library (libbia_hip.so) never links,
 * loads or calls it.
 *
 * All citations are file:line in the reference repository.
 */
#ifndef BIA_ORACLE_H
#define BIA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* branch codes reported per pair */
enum { ORC_HYPERBOLA = 0, ORC_PARABOLA = 1, ORC_ELLIPSE = 2, ORC_RKF = 3 };

/* flag bits reported per pair */
enum {
  ORC_F_STOCK_DEFAULT = 1, /* best-iterate slot never updated: reference returns its initial value */
  ORC_F_CYCLE = 2,         /* exact repeat of an earlier residual ended the Newton loop */
  ORC_F_ITR_CAP = 4,       /* iteration cap reached without |residual| < tolerance */
  ORC_F_ACOS_CLAMP = 8,    /* |cos(theta_ini)| > 1 clamp taken (hyperbola) */
  ORC_F_RKF_FAIL = 16,     /* per-pair RKF ended without reaching dt */
  ORC_F_NONFINITE = 32     /* a non-finite value in the output */
};

void orc_frame(const double R[3], const double G[3], double e_xi[3], double e_eta[3], double YZ[4]);
void orc_hyperbola(double eps, double dt, int itr, const double YZ[4], double YN[4], int *iters, int *flags);
void orc_ellipse(double eps, double dt, int itr, const double YZ[4], double YN[4], int *iters, int *flags);
void orc_parabola(double eps, double dt, const double YZ[4], double YN[4], int *iters, int *flags);
void orc_rkf_pair(double eps, double dt, const double YZ[4], double YN[4], int *nsteps, int *flags);
void orc_rkf_pair_tol(double eps, double dt, double relerr, double abserr, const double YZ[4], double YN[4],
                      int *nsteps, int *flags);

/* one pair: rel = (r_i - r_j, g_i - g_j); out = the six half-deltas credited to i */
void orc_pair(int method, double eps, double dt, int itr, const double rel[6], double out[6],
              int *branch, int *iters, int *flags);

/* pair kernel over explicit 2D inputs (batch): yz[4*n] -> yn[4*n], branch by classifier */
void orc_pairs2d(int method, double eps, double dt, int itr, int64_t n, const double *yz, double *yn,
                 int *branch, int *iters, int *flags);

/* whole BIA step over rows [i0, i1) (i1 = n for a whole step).  State and deltas are SoA
 * (x,y,z,u,v,w).  Deltas are accumulated in the reference's order, drift included.
 * hist (optional, 4 x 64 int64) counts pairs by branch and iteration count. */
void orc_bia_rows(int method, int n, double eps, double dt, int itr, const double *st,
                  int i0, int i1, double *delta, int64_t *hist);

/* invariants; out = KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za */
void orc_invariance_gpu(int n, double eps, const double *st, double out[11]);
void orc_invariance_cpu(int n, double eps, const double *st, double out[11]);
void orc_invariance_quad(int n, double eps, const double *st, double out[11]);

/* DIM direct-sum RKF45 over [*x0, xe] on an AoS state y[6*np].
 * step_t / step_y1 (optional, capacity cap) record each accepted step's end time and
 * particle-0 state.  Returns ier (0, 10000, 20000, 30000). */
int orc_dim_rkf45(int np, double eps, double *x0, double xe, double *y, double relerr, double abserr,
                  double *ese, int *nde, int *naccept, int *nreject, double *step_t, double *step_y1,
                  int cap);

/* coupled step: coupling 1 gravity with masses, 2 Coulomb with charges (and masses) */
void orc_bia_rows_coupled(int method, int n, int coupling, double coup, double dt, int itr, const double *st,
                          const double *mass, const double *charge, int i0, int i1, double *delta);

/* per-particle sums of the pair formulas' sensitivity to one rounding unit of their inputs */
int orc_conic_indeterminate(double eps, const double rel[6]);
void orc_sensitivity(int method, int n, int coupling, double eps, double dt, int itr, const double *st,
                     const double *mass, const double *charge, int samples, double *bpos, double *bvel);

/* two-body step in quadruple precision (closed-form Stumpff functions, bisection + Newton):
 * one pair's half deltas as hi + lo, and whole rows of selected particles of a system with the
 * scalar coupling (corr = sum of pair half deltas, delta = corr + drift), all [6][nrows] */
void orc_pair_quad(double eps, double dt, const double rel[6], double hi[6], double lo[6]);
void orc_rows_quad(int n, double eps, double dt, const double *st, int nrows, const int *rows, double *corr_hi,
                   double *corr_lo, double *delta);

/* per-pair RKF45 with the error test's tolerances as arguments, and whole rows of selected
 * particles with it (couplings as orc_bia_rows_coupled, eps for the scalar one) */
void orc_rows_rkf_tol(int n, int coupling, double coup, double eps, double dt, double relerr, double abserr,
                      const double *st, const double *mass, const double *charge, int nrows, const int *rows,
                      double *delta, int *failed);

/* one direct-sum derivative evaluation f = (v, a) on an AoS state */
void orc_dim_func(int np, double eps, const double *y, double *f);

#ifdef __cplusplus
}
#endif
#endif
