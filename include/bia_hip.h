/*
 * bia_hip.h -- C ABI of libbia_hip.so, the MI355X (gfx950) library for the pairwise hot path
 * of the Binary Interaction Approximation (BIA) N-body method.
 *
 * Two groups of entry points:
 *
 *  1. Reference-compatible symbols.  Same names, argument order, by-value/by-reference
 *     conventions and in-place semantics as the reference's CUDA objects, so the reference
 *     Fortran host programs (bia/bia.f90, dim/dim_3D.f90) link against this library
 *     unmodified.  They have no error return (as in the reference): on a HIP failure they
 *     print the cause and abort.
 *
 *  2. New entry points (prefix bia_).  They return an int status (BIA_OK = 0), never exit,
 *     and run on the context's stream.  Device-pointer variants (_dev) take state already
 *     resident in HBM in structure-of-arrays layout: six rows x[ld], y[ld], z[ld], u[ld],
 *     v[ld], w[ld] with row stride ld >= n.  Host variants (_host) copy in and out.
 *     A context serialises its own calls; use one context per host thread.  Its device
 *     workspace (tile partials, statistics slots, captured graphs) is shared by its calls, so
 *     work a call enqueued on one stream must finish before a call on another stream of the
 *     same context runs; calls on one stream are ordered by the stream.
 *
 * Units and conventions follow the reference: eps is the two-body coupling of the relative
 * equation of motion r'' = eps r/|r|^3 (eps = -2Gm for equal-mass gravity, bia/bia.f90:37),
 * dt the step, itr the Newton iteration cap (18 in bia/bia.f90:52).
 */
#ifndef BIA_HIP_H
#define BIA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BIA_OK 0
#define BIA_ERR_ARG 1     /* invalid argument */
#define BIA_ERR_HIP 2     /* HIP runtime error (text via bia_last_error) */
#define BIA_ERR_OOM 3     /* device allocation failed */
#define BIA_ERR_STATE 4   /* call sequence error (e.g. learned mode without trained weights) */
#define BIA_ERR_RKF 5     /* DIM RKF45 failure; the reference's IER code is returned separately */
#define BIA_ERR_IO 6      /* record file cannot be opened, read or written, or is malformed */

/* method codes (reference METHOD values, bia/bia.f90:135-139) */
#define BIA_METHOD_RKF 0  /* per-pair adaptive RKF45 */
#define BIA_METHOD_ALG 1  /* analytic ellipse / parabola / hyperbola */

/* solver modes of the analytic method (BIA_METHOD_ALG) */
#define BIA_MODE_PLAIN 0  /* the reference's conic solvers, starting guesses and iteration control */
#define BIA_MODE_LEARN 1  /* the Kepler drift; a pair whose series start is far off takes its second point from the learned table */
#define BIA_MODE_GUESS 2  /* the same Kepler drift, every pair from its series start (no network) */
#define BIA_MODE_PLAIN_LEARN 3  /* the reference's conic solvers and iteration control, the hyperbola and
                                   ellipse Newton started from network 1's anomaly; the reference's
                                   guess is the fallback and the judge (see bia_mlp_select) */

/* pair couplings of the batched entry points */
#define BIA_COUPLING_SCALAR 0  /* the reference's r'' = eps r/|r|^3, eps per system, equal-mass half split */
#define BIA_COUPLING_GRAVITY 1 /* eps_ij = -G (m_i + m_j), centre-of-mass split m_j/(m_i+m_j) : m_i/(m_i+m_j) */
#define BIA_COUPLING_COULOMB 2 /* eps_ij = k q_i q_j (1/m_i + 1/m_j), same split (unit masses if mass NULL) */

/* per-step statistics written by the step entry points (uint64 counters, BIA_STATS_LEN long) */
#define BIA_STATS_LEN 16
/* [0..3]  pairs per branch (hyperbola, parabola, ellipse, rkf); in the Kepler-drift modes the
 *         branch is the sign of the energy (W > 0, = 0, < 0)
 * [4..7]  summed solver iterations per branch (Newton evaluations / accepted RKF steps /
 *         evaluations of the universal time equation)
 * [8..15] pairs with status bit b set: 1 best-iterate never updated, 2 exact-repeat exit,
 *         4 iteration cap, 8 acos clamp, 16 RKF failure, 32 non-finite, 64 learned start used,
 *         128 learned start's run rejected for the reference guess's (BIA_MODE_PLAIN_LEARN) */

typedef struct bia_ctx bia_ctx;

/* ---------------------------------------------------------------------------------------- */
/* 1. reference-compatible symbols                                                          */
/* ---------------------------------------------------------------------------------------- */

/* Replaces find_dxyuv_kernel_ (bia/find_dXYUV_kernel.cu:526-561), called from
 * bia/bia.f90:135-139 as
 *   call find_dxyuv_kernel(%VAL(METHOD), %VAL(N), %VAL(I), %VAL(eps), %VAL(dt), %VAL(ITR),
 *                          dX, dY, dZ, dU, dV, dW, RXYZUVW, dXYZUVW)
 * I is 1-based.  RXYZUVW[6j..6j+5] = state_I - state_j (AoS).  Writes dXYZUVW[6j..] for j > I-1
 * and accumulates them into dX..dW: + at I-1, - at j.  The caller adds the drift. */
void find_dxyuv_kernel_(int METHOD, int N, int I, double EPS, double DT, int ITR, double *dX, double *dY,
                        double *dZ, double *dU, double *dV, double *dW, double *RXYZUVW, double *dXYZUVW);

/* Replaces invariance_gpu_ (bia/invariance.cu:50-110), called from bia/bia.f90:67-68.
 * KE = sum |v|^2, PE = eps * sum_{i<j} 1/r_ij, L = sum r x v, and the plain sums of x..w.
 * PE is accumulated in double-double (compensated) arithmetic. */
void invariance_gpu_(int N, double eps_th, double *X, double *Y, double *Z, double *U, double *V, double *W,
                     double *KE, double *PE, double *Lx, double *Ly, double *Lz, double *Ua, double *Va,
                     double *Wa, double *Xa, double *Ya, double *Za);

/* Replaces rkf_c_ (dim/rkf_kernel.cu:194-495), called from dim/dim_3D.f90:30.  Direct-sum
 * adaptive Fehlberg 4(5) with step doubling from *X0 to *XE on the AoS state YH0[neq]
 * (neq = 6 N).  On return YHN holds the state at *XE, *X0 = *XE, ESE = (YLN - YHN)/31.
 * H is in/out (initial step set when *INIT == 1).  YL0, YLN, ERR (neq) and WRK (11 neq)
 * are the caller's work arrays in the reference; the work lives on the device here and they
 * are left untouched.  On failure *IER = 10000 (too many
 * steps), 20000 (zero tolerance) or 30000 (step underflow) and, as in the reference, the
 * process exits with status 0 after printing the reason. */
void rkf_c_(int neq, double eps_th, double *X0, double *XE, double *YH0, int *INIT, const double *RELERR,
            const double *ABSERR, double *YHN, double *ESE, int *NDE, int *IER, double *H, double *YL0,
            double *YLN, double *ERR, double *WRK);

/* Replaces sincos_ (bia/sincos.c:3-6), the helper the reference CPU routines call. */
void sincos_(double *x, double *s, double *c);

/* ---------------------------------------------------------------------------------------- */
/* 2. new entry points                                                                      */
/* ---------------------------------------------------------------------------------------- */

int bia_version(void); /* major*10000 + minor*100 + patch */
const char *bia_error_string(int status);
const char *bia_last_error(bia_ctx *ctx);

int bia_create(bia_ctx **ctx, int device);
int bia_destroy(bia_ctx *ctx);
/* stream is a hipStream_t (0 = null stream); subsequent calls are ordered on it */
int bia_set_stream(bia_ctx *ctx, void *stream);
int bia_synchronize(bia_ctx *ctx);
/* Timing of the dominant kernel.  With profiling on, each step records HIP events on the
 * context's stream immediately around its pair kernel; bia_last_kernel_ms waits for the last
 * such pair and returns the elapsed milliseconds (BIA_ERR_STATE if none was recorded). */
int bia_set_profiling(bia_ctx *ctx, int on);
int bia_last_kernel_ms(bia_ctx *ctx, float *ms);
/* Durations in milliseconds of the last min(max, 64) pair-kernel launches recorded since
 * profiling was switched on, oldest first (waits for them); *n receives how many were written.
 * For timing steps enqueued back to back, with no synchronisation per step. */
int bia_kernel_ms(bia_ctx *ctx, float *ms, int max, int *n);
/* The same for the step's second kernel, the fixed-order reduction of the pair kernel's tile
 * partials into the deltas (an HBM-bound pass; its bytes per launch are the partials' size). */
int bia_reduce_kernel_ms(bia_ctx *ctx, float *ms, int max, int *n);

/* One whole BIA step of every unordered pair of an n-particle system.  No counterpart in the
 * reference ABI: it replaces the per-particle loop of bia/bia.f90:128-202 (N calls of
 * find_dxyuv_kernel_ with host scatter) by one device-resident pass.
 *   d_state  SoA [6][ld] positions and velocities (device)
 *   d_delta  SoA [6][ld] output: dX..dW including the drift v*dt (bia/bia.f90:194-196), i.e.
 *            the same values the reference accumulates into dX..dW (device)
 *   d_stats  BIA_STATS_LEN uint64 counters, accumulated into (device, may be NULL) */
int bia_step_dev(bia_ctx *ctx, int method, int mode, int n, double eps, double dt, int itr,
                 const double *d_state, int ld, double *d_delta, unsigned long long *d_stats);

/* The step of bia_step_dev (same kernels, bitwise the same deltas) that allocates nothing.  No
 * counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).  Its plan and partial
 * sums live in a library-owned static device array of 31.0 MiB that each GPU the library runs
 * on holds until the process ends, which limits n to 1 <= n <= 8192 (BIA_ERR_ARG beyond,
 * d_delta untouched).  No statistics and no failure records; the context's autotuned grids do
 * not apply.  Every call on one GPU shares that GPU's array (from any context): calls that
 * overlap on one GPU race, so order them on one stream. */
int bia_step_noalloc_dev(bia_ctx *ctx, int method, int mode, int n, double eps, double dt, int itr,
                         const double *d_state, int ld, double *d_delta);

/* Batched step: nsys independent systems of uneven sizes in one persistent launch.  The systems
 * are concatenated column-wise in one SoA state [6][ld] (system s occupies the next sizes[s]
 * columns); d_delta has the same layout.  No counterpart in the reference ABI (its only ensemble
 * is the sequential N_calc loop of bia/bia.f90:62).
 *   coupling   BIA_COUPLING_*; eps (host, nsys) for SCALAR, coup = G or k for the others
 *   d_mass, d_charge  per-particle arrays (device, ld) for GRAVITY / COULOMB, else NULL */
int bia_step_batched_dev(bia_ctx *ctx, int method, int mode, int coupling, int nsys, const int *sizes,
                         const double *eps, double coup, double dt, int itr, const double *d_state, int ld,
                         const double *d_mass, const double *d_charge, double *d_delta, unsigned long long *d_stats);

/* Autotuner: times the pair kernel's persistent grid size and register-occupancy variant on this
 * input (one step per candidate, written to d_delta) and keeps the fastest for this list of sizes
 * in the context.  *best_grid receives grid * 10 + occupancy variant (1 default, 4 or 5 waves/SIMD). */
int bia_autotune_dev(bia_ctx *ctx, int method, int mode, int coupling, int nsys, const int *sizes, const double *eps,
                     double coup, double dt, int itr, const double *d_state, int ld, const double *d_mass,
                     const double *d_charge, double *d_delta, float *best_ms, int *best_grid);

/* Device-resident multi-step driver: nsteps of (step, state += delta) enqueued on the context's
 * stream with no host synchronisation; with use_graph the step is captured once as a hipGraph and
 * replayed.  d_state is updated in place; d_delta holds the last step's deltas.  d_stats (device,
 * BIA_STATS_LEN, may be NULL) accumulates the per-pair statistics of every step. */
int bia_run_steps_dev(bia_ctx *ctx, int method, int mode, int coupling, int nsys, const int *sizes, const double *eps,
                      double coup, double dt, int itr, double *d_state, int ld, const double *d_mass,
                      const double *d_charge, double *d_delta, int nsteps, int use_graph,
                      unsigned long long *d_stats);

/* Workspace of the step.  The pair kernel leaves partial sums that a second kernel adds in a
 * fixed order (bitwise reproducible results): per block of four 64-particle row tiles, one column
 * partial per later particle and one row partial set per work item of up to four column tiles;
 * for one system of n particles about 12 x 8 x n x n_pad / 256 bytes with four-tile items (n_pad =
 * 64 ceil(n / 64); 0.81 GB at n = 65536), up to four times that for the single-tile items of
 * small systems (1 MB at n = 1332).  When that exceeds the context's limit (default 1 GiB; 0 = no
 * limit), a single-system step runs in stages of G blocks, each followed by its reduction (G the
 * largest whose partials fit, at least 1): the same deltas bitwise, in O(n) workspace.  Batched
 * steps of several systems are not staged.  No counterpart in the reference (its per-particle
 * calls keep 96 n bytes on the device, bia/find_dXYUV_kernel.cu:532-533). */
int bia_set_workspace_limit(bia_ctx *ctx, unsigned long long bytes);
/* The plan of the context's last step: its number of stages and tile-partial workspace in bytes
 * (BIA_ERR_STATE before any step). */
int bia_plan_info(bia_ctx *ctx, int *nstages, unsigned long long *workspace_bytes);

/* Tolerances of the per-pair RKF45 (BIA_METHOD_RKF) for the context's subsequent steps: the error
 * test of a trial step is |err_k| <= relerr (|y_k| + |y_k'|)/2 + abserr per component, as in
 * bia/bia-rkf.cu:72-79, where RELERR = 0 and ABSERR = 1e-16 are compile-time constants (bia-rkf.cu:5;
 * they stay the defaults).  Both >= 0, finite, not both 0; else BIA_ERR_ARG.  The per-particle
 * reference symbol find_dxyuv_kernel_ keeps the reference's values. */
int bia_set_rkf_tolerance(bia_ctx *ctx, double relerr, double abserr);

/* Per-pair failure records.  The reference's CPU pair loop writes every pair whose solver did not
 * converge to its units 88 (RKF) and 99 (ALG) with (realisation, i, j, IER, iterations, error)
 * (bia/bia.f90:151-154,167-170); its GPU path caps such pairs silently.  With a log set, every
 * subsequent step of the context writes one bia_pair_record per pair that stopped at its
 * iteration cap, whose RKF45 ended before dt, or whose output is not finite, into d_records
 * (device, capacity records) and adds one to *d_count (device, the caller zeroes it) per such
 * pair, including those beyond the capacity, which are not written.  Records are in no particular
 * order (sort by sys, i, j).  Such steps run the pair kernel's statistics variant.  d_count NULL
 * switches logging off.  The buffers are the caller's and must outlive the steps that use them. */
typedef struct {
  int sys;        /* system of the call's list */
  int i, j;       /* the pair, 0-based within the system, i < j */
  int branch;     /* 0 hyperbola, 1 parabola, 2 ellipse, 3 RKF45 (Kepler drift: sign of the energy) */
  int iters;      /* evaluations / accepted RKF steps */
  unsigned flags; /* status bits of the statistics (4 iteration cap, 16 RKF failure, 32 non-finite, ...) */
  double error;   /* last residual: the reference solvers' best residual of their dimensionless time
                     equation; the Kepler drift's t(s) - dt; the RKF45's error ratio of its last trial */
} bia_pair_record;
int bia_set_failure_log(bia_ctx *ctx, bia_pair_record *d_records, int capacity, unsigned long long *d_count);

/* Host-array variant of bia_step_dev for Fortran callers (bind(C) interface in
 * bia_3d_gpu_amd/fortran/bia_hip.f90).  dX..dW are overwritten; stats (may be NULL) is
 * BIA_STATS_LEN counters for this call. */
int bia_step_host(bia_ctx *ctx, int method, int mode, int n, double eps, double dt, int itr, const double *X,
                  const double *Y, const double *Z, const double *U, const double *V, const double *W,
                  double *dX, double *dY, double *dZ, double *dU, double *dV, double *dW,
                  unsigned long long *stats);

/* Invariants of an SoA state; out[11] = KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za
 * (the order of invariance_gpu_'s outputs).  d_out is device memory. */
int bia_invariants_dev(bia_ctx *ctx, int n, double eps, const double *d_state, int ld, double *d_out);
int bia_invariants_host(bia_ctx *ctx, int n, double eps, const double *X, const double *Y, const double *Z,
                        const double *U, const double *V, const double *W, double *out);
/* The same for nsys systems side by side in one [6, ld] device state (columns in the order of
 * sizes, eps per system; both host arrays): d_out[11 s + k] is invariant k of system s.  One
 * launch of each kernel for the whole list; each system's values are bitwise those of
 * bia_invariants_dev on its columns alone.  No counterpart in the reference (its invariance is per
 * realisation, bia/bia.f90:236-241): for the monitored multi-step driver over ensembles. */
int bia_invariants_batched_dev(bia_ctx *ctx, int nsys, const int *sizes, const double *eps, const double *d_state,
                               int ld, double *d_out);

/* The invariants of bia_invariants_batched_dev with the couplings of the batched step: KE, L and
 * the position and velocity sums weighted by the masses (unit masses for COULOMB with d_mass
 * NULL), and PE = -2 G sum_{i<j} m_i m_j / r (GRAVITY, coup = G) or 2 k sum_{i<j} q_i q_j / r
 * (COULOMB, coup = k), so that KE + PE is twice the total energy, as KE + PE of the reference's
 * equal-mass form is 2 E / m.  d_mass, d_charge: device arrays with the state's columns.  No
 * counterpart in the reference (its invariants are for equal masses, bia/invariance.cu:15-48):
 * the energy check of ensembles with a mass spectrum or charges. */
int bia_invariants_coupled_batched_dev(bia_ctx *ctx, int coupling, double coup, int nsys, const int *sizes,
                                       const double *d_state, int ld, const double *d_mass, const double *d_charge,
                                       double *d_out);

/* DIM direct-sum RKF45 over [*x0, xe] on an AoS state y[6 np] (host, in/out), with the
 * reference's step controller (dim/rkf_kernel.cu:237-491).  *h is the step (in/out; set
 * when *init == 1).  ese (may be NULL) receives (YLN - YHN)/31.  Counters may be NULL.
 * Returns BIA_OK, or BIA_ERR_RKF with *ier = 10000/20000/30000. */
int bia_dim_rkf45_host(bia_ctx *ctx, int np, double eps, double *x0, double xe, double *y, int *init,
                       double relerr, double abserr, double *h, double *ese, int *nde, int *naccept,
                       int *nreject, int *ier);

/* The same DIM integration with the pair couplings of the batched step: the direct-sum
 * acceleration of particle i is
 *   BIA_COUPLING_GRAVITY  a_i = -G sum_j m_j (r_i - r_j) / |r_i - r_j|^3            (coup = G)
 *   BIA_COUPLING_COULOMB  a_i = (k q_i / m_i) sum_j q_j (r_i - r_j) / |r_i - r_j|^3  (coup = k,
 *                         unit masses if mass is NULL)
 * so that a pair's relative motion has the eps_ij of the step's coupling.  mass, charge: host
 * arrays (np).  No counterpart in the reference, whose DIM has one eps for all pairs
 * (dim/rkf_kernel.cu:29-53): the accuracy reference for ensembles with a mass spectrum or
 * charges.  Returns BIA_ERR_ARG for another coupling or a missing array. */
int bia_dim_rkf45_coupled_host(bia_ctx *ctx, int coupling, double coup, int np, const double *mass,
                               const double *charge, double *x0, double xe, double *y, int *init, double relerr,
                               double abserr, double *h, double *ese, int *nde, int *naccept, int *nreject,
                               int *ier);

/* ---------------------------------------------------------------------------------------- */
/* learned initialiser (no counterpart in the reference)                                     */
/* ---------------------------------------------------------------------------------------- */
/* A 3 -> 64 -> 64 -> 1 GELU network proposes the starting universal anomaly of the Kepler drift
 * (BIA_MODE_LEARN): a pair whose first Halley correction from the rational series start exceeds a
 * tenth of the anomaly (the start was far off) takes the table's proposal as its second point,
 * inside the bracket the first evaluation set; every other pair runs as in BIA_MODE_GUESS.
 * Inputs: the squashed dimensionless
 * numbers u_k = sign(a_k) log2(1 + |a_k| / 1e-2) / log2(1 + 1e4) (single precision) of
 * a1 = eta dt / r0^2, a2 = zeta dt^2 / r0^3,
 * a3 = beta dt^2 / r0^2 (eta = r.v, beta = 2 mu / r0 - v^2, zeta = mu - beta r0, mu = -eps);
 * output log(s / s_start) / min(|a|, 1)^3, the converged anomaly relative to the rational
 * series start.  Sample rows are float [u1, u2, u3, target].  Weights are 4481
 * floats: W1[64][3] b1[64] W2[64][64] b2[64] w3[64] b3.  Products run on BF16 MFMA with FP32
 * accumulation; the master weights are FP32. */
int bia_mlp_nparam(void);
/* A context holds two such networks, with their own weights, optimiser state and table:
 *   net 0 (default): the Kepler drift's start above (BIA_MODE_LEARN);
 *   net 1: the starting anomaly of the reference's own Newton iterations (BIA_MODE_PLAIN_LEARN),
 *          replacing the hyperbola's theta_ini / theta_gss pair (bia/find_dXYUV_kernel.cu:169-185)
 *          and the ellipse's xi_2 guess (:253-265).  Inputs, u3 = +1 hyperbola / -1 ellipse:
 *          hyperbola u1 = log10(e - 1) / 7 (clamped to [-1, 1]), u2 = s (log2|T| + 10) / 37
 *          (magnitude clamped to [0, 1]) of the dimensionless target time T, s = -1 attractive / +1 repulsive
 *          (the sign of cos theta_0), target e log(theta* / theta_gss); ellipse
 *          u1 = 1 - 2 sqrt(1 - e), u2 = M_r / pi of the end mean anomaly reduced to [-pi, pi],
 *          target sin E* at the root of Kepler's equation.  The pair kernel reads the table's end slices u3 = -1,
 *          +1.  The learned start's iteration keeps the reference's control (best iterate, exact
 *          repeat, cap); where it stops at the cap unconverged, the reference's own guess runs as
 *          well and the smaller best residual is kept.
 * bia_mlp_select(ctx, net) makes `net` the one the bia_mlp_* and bia_record_samples_dev calls
 * below act on (0 or 1; BIA_ERR_ARG otherwise); the step entry points read each mode's own
 * network whichever is selected. */
int bia_mlp_select(bia_ctx *ctx, int net);
int bia_mlp_init(bia_ctx *ctx, unsigned long long seed);
int bia_mlp_set_weights(bia_ctx *ctx, const float *w, int n);
int bia_mlp_get_weights(bia_ctx *ctx, float *w, int n);
/* d_rows [nrows][4] (device; the 4th column is ignored) -> d_y [nrows] (device) */
int bia_mlp_forward_dev(bia_ctx *ctx, const float *d_rows, int nrows, float *d_y);
/* gradient of mean((y - t)^2) over all rows; d_grad (device) receives 4481 floats + the loss */
int bia_mlp_grad_dev(bia_ctx *ctx, const float *d_rows, int nrows, float *d_grad);
/* Adam, cosine-decayed learning rate, random mini-batches of `batch` rows per step;
 * loss_curve (host, may be NULL) receives the loss of each step */
int bia_mlp_train_dev(bia_ctx *ctx, const float *d_rows, int nrows, int steps, int batch, float lr,
                      unsigned long long seed, float *loss_curve);
/* tabulate the network on an n x n x n grid of (u1, u2, u3) in [-1, 1]^3 (2 <= n <= 256) for the
 * pair kernel's trilinear lookup; enables BIA_MODE_LEARN */
int bia_mlp_build_table(bia_ctx *ctx, int n);
/* record nsamples pairs of a system that are not weak over dt (each of up to 4096 random draws):
 * rows [u1, u2, u3, target] solved to convergence, u1 = NaN when none was found (device).
 * With net 1 selected: rows of the reference solvers' starts (above) from hyperbolic pairs
 * (even samples) and elliptic pairs (odd samples) that the reference's iteration solves to its
 * exit without the cap (hyperbola) or whose root of Kepler's equation bisection finds (ellipse). */
int bia_record_samples_dev(bia_ctx *ctx, int n, double eps, double dt, const double *d_state, int ld, int nsamples,
                           unsigned long long seed, float *d_out);

/* ---- the reference's record files (data plane; host only, no context) ----
 * Fortran unformatted sequential files with 4-byte little-endian record markers, as the
 * reference's programs read and write them: .bbb, one record of 6N doubles per realisation,
 * particle-major x, y, z, u, v, w (IC/bbb_maker_REV.f90:71; the N_calc loop of bia/bia.f90:62-64
 * reads one per pass; nhead = 0); .DIM, records [t, 6N] (dim/dim_3D.f90:32; nhead = 1); .BIA,
 * records [dt, 6N deltas] (bia/bia.f90:206; nhead = 1).  Arrays on the caller's side are host
 * memory in the step entry points' layout (six rows of stride ld).  No counterpart in the
 * reference's C ABI: its Fortran reads and writes these files itself; these are for C and C++
 * callers.  BIA_ERR_IO: the file cannot be opened, read or written, or a record is malformed (a
 * marker that is negative, does not fit the file or differs from its closing marker, or a payload
 * that is not nhead doubles and whole particles).
 * bia_records_scan: number of records and the largest particle count among them.
 * bia_record_read: record `rec` (0-based; BIA_ERR_ARG past the last) into head[nhead] and
 *   soa[6][ld] (ld >= its particle count); with soa NULL only *n is returned.
 * bia_record_write: one record of n particles, truncating the file (append 0) or appending (1);
 *   at most (2^31 - 1 - 8 nhead) / 48 particles (the marker's limit). */
int bia_records_scan(const char *path, int nhead, long long *nrec, long long *nmax);
int bia_record_read(const char *path, int nhead, long long rec, double *head, double *soa, long long ld,
                    long long *n);
int bia_record_write(const char *path, int append, int nhead, const double *head, const double *soa, long long ld,
                     long long n);

#ifdef __cplusplus
}
#endif
#endif /* BIA_HIP_H */
