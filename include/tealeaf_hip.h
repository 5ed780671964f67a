/*
 * tealeaf_hip.h -- C ABI of libtealeaf_hip, the MI355X (gfx950) linear-solve
 * hot path of the TeaLeaf 2-D heat-conduction mini-app.
 *
 * The ABI stands in for the solver plug-in interface of Laura7089/TeaLeaf.jl:
 * a solver is a module whose `solve!(chunk::Chunk, set::Settings, rx, ry)` is
 * called once per timestep (src/TeaLeaf.jl:74) and selected by `use_*` deck
 * lines (src/settings.jl:103-111).  Every entry point returns an int status
 * (TL_OK = 0); no C++ exception crosses the ABI.  A context owns one HIP
 * stream on one device and is not thread-safe.
 *
 * Host arrays are Float64, column-major with x fastest, of size
 * (nx + 2*halo_depth) * (ny + 2*halo_depth) -- the layout of the reference's
 * Chunk fields (src/chunk.jl:19-38, 68-70).
 */
#ifndef TEALEAF_HIP_H
#define TEALEAF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the reference throws strings instead (CG.jl:49, kernels.jl:42,80) */
#define TL_OK 0
#define TL_E_INVALID 1      /* invalid argument or deck value, e.g. coefficient (CG.jl:48-50) */
#define TL_E_HIP 2          /* HIP runtime failure */
#define TL_E_NOMEM 3        /* device allocation failed */
#define TL_E_BREAKDOWN 4    /* p.Ap <= 0 or a non-finite residual */
#define TL_E_EIGEN 5        /* negative eigenvalue or eigen-solve failure (kernels.jl:41-43, 80) */
#define TL_E_PARSE 6        /* malformed deck */
#define TL_E_STATE 7        /* call out of order (e.g. a learned warm start without a network) */
#define TL_E_IO 8           /* file could not be read (checkingvalue's open, settings.jl:181) */
#define TL_E_CHECK 9        /* the solve claimed convergence but the FP64 residual of the returned u,
                               recomputed by separate kernels, does not meet the rule (tl_check_solution) */

/* solvers, settings.jl:103-111 */
#define TL_CG 0
#define TL_CHEBY 1
#define TL_PPCG 2
#define TL_JACOBI 3

/* arithmetic of the solve */
#define TL_FP64 0           /* everything in FP64 (the reference's Float64) */
#define TL_FP32_IR 1        /* FP32 inner CG (FP32 vectors and arithmetic, the cell coefficient stored
                               in 16 bits: see TL_COEF_*); FP64 residual, solution and reduction
                               accumulators */

/* stop rules */
#define TL_STOP_REFERENCE 0 /* each solver's own rule: CG sqrt|rr| < eps (CG.jl:24); others |err| < eps */
#define TL_STOP_RELATIVE 1  /* CG: ||u0 - A u||_2 <= rel_tol * ||u0||_2, checked on the true FP64 residual */

/* initial guess of the solve */
#define TL_WARM_NONE 0      /* u = energy*density, the reference's start (CG.jl:52) */
#define TL_WARM_LEARNED 1   /* U-Net prediction; falls back to TL_WARM_NONE when its residual is worse */
#define TL_WARM_GUESS 2     /* caller-provided guess set with tl_set_guess, same fallback */

/* tl_solve_params.launch: how the CG iterations of TL_CG solves (FP64 and
 * TL_FP32_IR) reach the GPU.  Both give bit-identical results.  Other solvers
 * always use stream launches. */
#define TL_LAUNCH_STREAM 0  /* kernels enqueued in batches of up to 64 iterations */
#define TL_LAUNCH_GRAPH 1   /* after a first stream batch, replays of one hipGraph of 64 iterations
                               (captured on first use, kept with the context, re-captured when
                               a buffer or coefficient it holds changes); tl_profile_* does not
                               see the replayed kernels */

/* field ids, in the order of the Chunk struct (chunk.jl:25-38) */
#define TL_F_DENSITY0 0
#define TL_F_DENSITY 1
#define TL_F_ENERGY0 2
#define TL_F_ENERGY 3
#define TL_F_U 4
#define TL_F_U0 5
#define TL_F_P 6
#define TL_F_R 7
#define TL_F_MI 8
#define TL_F_W 9
#define TL_F_KX 10
#define TL_F_KY 11
#define TL_F_SD 12
#define TL_NFIELDS 13

/* halo exchange mask, EXCHANGE_FIELDS of chunk.jl:12 */
#define TL_X_DENSITY (1u << 0)
#define TL_X_P (1u << 1)
#define TL_X_ENERGY0 (1u << 2)
#define TL_X_ENERGY (1u << 3)
#define TL_X_U (1u << 4)
#define TL_X_SD (1u << 5)

typedef struct tl_ctx tl_ctx;

/* solver settings of Settings (settings.jl:39-78) plus the build's additions */
typedef struct {
    int solver;           /* TL_CG ... TL_JACOBI            (Settings.solver) */
    int coefficient;      /* 1 conductivity, 2 reciprocal    (Settings.coefficient) */
    int max_iters;        /*                                 (Settings.maxiters) */
    int presteps;         /* CG steps before a Cheby/PPCG switch (Settings.presteps) */
    int ppcg_inner_steps; /*                                 (Settings.ppcginnersteps) */
    int error_switch;     /*                                 (Settings.errorswitch) */
    double eps;           /*                                 (Settings.eps) */
    double epslim;        /*                                 (Settings.epslim) */
    int precision;        /* TL_FP64 | TL_FP32_IR */
    int stop_rule;        /* TL_STOP_REFERENCE | TL_STOP_RELATIVE */
    double rel_tol;       /* for TL_STOP_RELATIVE */
    int warm_start;       /* TL_WARM_* */
    int check_result;     /* Settings.checkresult: residual! in tl_solve_finished */
    double ir_delta;      /* TL_FP32_IR: refresh the FP64 residual when the FP32 one falls by this factor
                             (0 < ir_delta < 1; any other value selects the default 0.1) */
    int launch;           /* TL_LAUNCH_* */
    int fp32_kernels;     /* TL_FP32_IR: TL_FP32_KERNELS_* */
    int warm_policy;      /* TL_WARM_LEARNED: TL_WARM_POLICY_* */
    int ppcg_pairs;       /* TL_PPCG: TL_PPCG_PAIRS_* */
    int cheby_pairs;      /* TL_CHEBY: TL_CHEBY_PAIRS_* */
    int reserved[3];
} tl_solve_params;

/* tl_solve_params.warm_policy: when a TL_WARM_LEARNED solve runs the network.
 * Either way an attempted prediction is kept only if its residual beats the
 * previous-step start's. */
#define TL_WARM_POLICY_BACKOFF 0 /* on the context's first learned solve and after every accepted
                                    prediction; after a rejection the next attempt waits 2, 4, 8,
                                    then 16 solves (a non-attempted solve starts from u) */
#define TL_WARM_POLICY_EVERY 1   /* every solve */

/* tl_solve_params.ppcg_pairs: PPCG's inner Chebyshev steps (PPCG.jl:78-83)
 * two per sweep over the grid, or one.  Bit-identical results. */
#define TL_PPCG_PAIRS_AUTO 0 /* two steps per sweep when ppcg_inner_steps >= 4 (an odd last step alone) */
#define TL_PPCG_PAIRS_OFF 1  /* one step per sweep */

/* tl_solve_params.cheby_pairs: Chebyshev iterations (Cheby.jl:93-107) two per
 * sweep over the grid, or one.  Bit-identical results. */
#define TL_CHEBY_PAIRS_AUTO 0 /* two iterations per sweep where the first needs no residual norm */
#define TL_CHEBY_PAIRS_OFF 1  /* one iteration per sweep */

/* tl_solve_params.fp32_kernels: how a TL_FP32_IR iteration is split into kernels.
 * Both reach the same FP64 tolerance; their FP32 roundings differ. */
#define TL_FP32_KERNELS_AUTO 0   /* nx a multiple of 4: pairs with stream launches, fused with
                                   TL_LAUNCH_GRAPH; else two kernels */
#define TL_FP32_KERNELS_TWO 1    /* two kernels and two reductions per iteration (34 B/cell) */
#define TL_FP32_KERNELS_FUSED 2  /* one kernel and one reduction per iteration (22 B/cell); needs nx % 4 == 0 */
#define TL_FP32_KERNELS_PAIRS 3  /* two iterations per two kernels and two reductions (s-step CG, s = 2;
                                   18 B/cell per iteration); needs nx % 4 == 0; stream launches */

/* replaces the reference's @info lines (CG.jl:27, Cheby.jl:60, PPCG.jl:54, Jacobi.jl:29) */
typedef struct {
    int iterations;       /* outer iterations (tt at exit) */
    int cg_iters;         /* CG iterations (before the switch for Cheby/PPCG) */
    int cheby_iters;      /* Chebyshev / PPCG iterations after the switch */
    int est_iters;        /* Chebyshev iteration estimate (Cheby.calciter) */
    int inner_iters;      /* TL_FP32_IR: FP32 CG iterations */
    int refreshes;        /* TL_FP32_IR: FP64 residual recomputations; FP64 CG with TL_STOP_RELATIVE:
                             restarts from the true residual when the recursive one met the rule
                             but the true one did not */
    int converged;        /* stop rule met (0: ran max_iters) */
    int fallback;         /* warm start rejected in favour of u = energy*density */
    double error;         /* final error as the solver defines it */
    double eigmin, eigmax;
    double rel_resid;     /* ||u0 - A u|| / ||u0|| of the returned u, FP64 */
    double init_rel_resid;/* same for the initial guess that was used */
    double solve_ms;      /* wall time of tl_solve */
    int start;            /* initial guess used: TL_START_* */
    int coef_format;      /* TL_FP32_IR: storage of the inner CG's cell coefficient, TL_COEF_*; 0 otherwise */
    int warm_attempted;   /* a warm start was evaluated in this solve (network run or guess tested) */
    int reserved_i;
    double warm_ms;       /* wall time of that evaluation (prediction and residual test) */
    double reserved[1];
} tl_solve_stats;

/* tl_solve_stats.coef_format: half with a power-of-two scale when the coefficient's range
 * fits half's normal range (max/min <= 2^28), else bfloat16 */
#define TL_COEF_FP16 1
#define TL_COEF_BF16 2
#define TL_COEF_FP32 3

/* tl_solve_stats.start */
#define TL_START_PREVIOUS 0 /* u = energy*density (the reference's start) */
#define TL_START_WARM 1     /* the learned prediction or the caller's guess */

/* a parsed tea.in deck (settings.jl:90-173); states keep the deck's coordinates */
#define TL_MAX_STATES 128
typedef struct {
    double density, energy, xmin, ymin, xmax, ymax, radius;
    int geometry;         /* 0 rectangle, 1 circular, 2 point */
    int defined;
} tl_state;

typedef struct {
    int x_cells, y_cells, halo_depth;
    double xmin, ymin, xmax, ymax, dx, dy;
    long long end_step;
    int summary_frequency;
    double dt_init, end_time;
    int check_result;
    tl_solve_params params;
    int nstates;
    tl_state states[TL_MAX_STATES];
    int n_warnings;       /* unknown keys, as settings.jl:109,127 warn */
} tl_deck;

/* ------------------------------------------------------------ lifecycle */
void tl_default_params(tl_solve_params *p);
/* Chunk(settings), chunk.jl:68-89: nx*ny interior cells, domain origin and cell size;
 * TL_E_INVALID unless nx, ny >= 2 and 2 <= halo_depth <= 16 (with a depth-1 halo the
 * reference's CG.init! divides by an unset coefficient, CG.jl:59-68) */
int tl_create(tl_ctx **ctx, int nx, int ny, int halo_depth, double xmin, double ymin, double dx, double dy,
              int device);
void tl_destroy(tl_ctx *ctx);
const char *tl_last_error(const tl_ctx *ctx);
int tl_dims(const tl_ctx *ctx, int *X, int *Y);

/* ---------------------------------------------------------------- fields */
/* H2D / D2H of one reference field in the host layout described above */
int tl_set_field(tl_ctx *ctx, int field_id, const double *host);
int tl_get_field(tl_ctx *ctx, int field_id, double *host);
/* density and energy as the Julia driver holds them before solve! (TeaLeaf.jl:71-74) */
int tl_set_fields(tl_ctx *ctx, const double *density, const double *energy);
/* caller's initial guess (host, full layout) for TL_WARM_GUESS */
int tl_set_guess(tl_ctx *ctx, const double *host);

/* ----------------------------------------------------- per-step operations */
/* setchunkstate! + initialiseapp!, chunk.jl:122-151 and TeaLeaf.jl:35-44, on the device */
int tl_initialise(tl_ctx *ctx, const tl_state *states, int nstates);
/* haloupdate!, kernels.jl:146-159 (reflective, depth layers, interior span);
 * depth <= min(halo_depth, nx, ny) */
int tl_halo_update(tl_ctx *ctx, unsigned field_mask, int depth);
/* set.solver.solve!(chunk, set, rx, ry), TeaLeaf.jl:74 */
int tl_solve(tl_ctx *ctx, const tl_solve_params *p, double rx, double ry, tl_solve_stats *out);
/* solvefinished!, kernels.jl:166-170: residual! (if check_result), finalise!, halo(energy) */
int tl_solve_finished(tl_ctx *ctx, int check_result);
/* fieldsummary, kernels.jl:119-133: out = {vol, mass, ie, temp} */
int tl_field_summary(tl_ctx *ctx, double out[4]);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc): tl_field_summary of caller-owned DEVICE fields in the host
 * layout ((nx + 2*halo_depth) * (ny + 2*halo_depth) doubles, x fastest),
 * enqueued on `stream` (a hipStream_t; NULL = the null stream), writing
 * {vol, mass, ie, temp} to out_dev (4 doubles of device memory).  No context
 * and no allocation: the per-block partial sums live in a static device array
 * of the library (256 KiB per GPU, held until the process ends), so the grid
 * may have at most 8192 summation blocks (256-column strips x row bands of up
 * to 1024 rows: any grid up to 32768^2); larger grids return TL_E_INVALID.
 * The same kernels and summation order as tl_field_summary, so the results
 * are bit-identical.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU (other streams, other threads) race. */
int tl_field_summary_noalloc(const double *density, const double *energy, const double *u, int nx, int ny,
                             int halo_depth, double dx, double dy, double *out_dev, void *stream);
/* one timestep of diffuse!, TeaLeaf.jl:67-81 (dt = Settings.dtinit) */
int tl_step(tl_ctx *ctx, const tl_solve_params *p, double dt, tl_solve_stats *out);

/* -------------------------------------------------------------- deck / run */
int tl_deck_parse(const char *text, tl_deck *out);
/* Chunk(settings) + initialiseapp! for a parsed deck */
int tl_deck_create(tl_ctx **ctx, const tl_deck *deck, int device);
/* whole diffuse! loop (TeaLeaf.jl:62-83) on the device; cb (may be NULL) sees every step */
typedef void (*tl_step_cb)(int step, const double summary[4], const tl_solve_stats *stats, void *user);
int tl_run(tl_ctx *ctx, const tl_deck *deck, const tl_solve_params *p, int nsteps, tl_step_cb cb, void *user);

/* checkingvalue, settings.jl:180-196: the 4th field of the first line
 * "xcells ycells endstep value" of problemfile whose three integers match;
 * *value = 1.0 and *found = 0 when no line matches (the reference's fallback).
 * TL_E_IO if the file cannot be opened, TL_E_PARSE for a malformed line. */
int tl_checking_value(const char *problemfile, int xcells, int ycells, int endstep, double *value, int *found);
/* fieldsummary's QA rule, kernels.jl:127-132: *qa_diff = |100 actual/cv - 100|;
 * returns 1 (PASSED) if *qa_diff < 0.001, else 0 (FAILED) */
int tl_qa_check(double actual, double checking_value, double *qa_diff);

/* ------------------------------------------------------ host-side scalars */
/* eigenvalues! (kernels.jl:19-51): extrema of the Lanczos tridiagonal of n CG
 * steps (alpha, beta 0-based, n entries), scaled by 0.95 / 1.05 */
int tl_eigen_bounds(const double *alpha, const double *beta, int n, double *eigmin, double *eigmax);
/* Cheby.coef!, Cheby.jl:121-135 (alpha, beta 0-based, n entries) */
void tl_cheby_coef(double eigmin, double eigmax, int n, double *alpha, double *beta, double *theta);
/* Cheby.calciter, Cheby.jl:109-118 */
int tl_cheby_calciter(double eigmin, double eigmax, double error, double bb);
/* The check tl_solve applies to every solve before it returns (no counterpart
 * in the reference, which trusts its recursive residual, CG.jl:24).  Inputs:
 * the solve's parameters and statistics (st->iterations, st->converged), bb =
 * sum u0^2 as the solve itself reduced it (negative: the solver does not form
 * it, Jacobi), and two FP64 sums recomputed from the returned u by separate
 * kernels, u0sq = sum u0^2 and resid2 = sum (u0 - A u)^2 over the interior.
 * Returns TL_OK when the claim holds, else TL_E_CHECK:
 *   - bb >= 0 and u0sq disagree by more than 1e-9 relative (one of the two
 *     reductions did not see its inputs);
 *   - a CG solve (either precision) under TL_STOP_RELATIVE claims
 *     convergence but sqrt(resid2) > rel_tol sqrt(u0sq) (1 + 1e-6; the slack
 *     covers the summation order of the two residual kernels);
 *   - a CG-family solve (CG, or the CG phase of Chebyshev / PPCG) claims an
 *     exact start (converged after 0 iterations) while the recomputed
 *     residual misses the rule's threshold (sqrt(resid2) < eps for the
 *     reference rule, as above for the relative one). */
int tl_check_solution(const tl_solve_params *p, const tl_solve_stats *st, double bb, double u0sq, double resid2);

/* ------------------------------------------------- device-side diagnostics */
/* w = A*field over the interior (smvp, kernels.jl:178-184), result left in TL_F_W */
int tl_apply_operator(tl_ctx *ctx, int field_id, double rx, double ry, int coefficient);
/* CG history (cgα, cgβ of chunk.jl:56-57) of the last solve, n entries */
int tl_get_cg_history(tl_ctx *ctx, double *alpha, double *beta, int n);

/* ------------------------------------------------------------ profiling */
/* kernel ids for tl_profile_begin.  A CG iteration is two kernels:
 * PW = end of the previous iteration (beta, stop test, p = r + beta p) fused
 * with w = A p and <p, w>; UR = u += alpha p, r -= alpha w and <r, r>. */
#define TL_K_CG_W64 0  /* FP64 PW (also the unfused w = A p of Chebyshev/PPCG's CG presteps) */
#define TL_K_CG_UR64 1 /* FP64 UR */
#define TL_K_CG_P64 2  /* FP64 p = r + beta p: a solve's last iteration, and the CG presteps */
#define TL_K_CG_W32 3  /* FP32 PW of the inner CG (w = A p only reduced, not stored) */
#define TL_K_CG_UR32 4 /* FP32 UR of the inner CG (recomputes A p) */
#define TL_K_CG_P32 5  /* unused: folded into TL_K_CG_W32 */
#define TL_K_CHEBY64 6 /* Chebyshev iteration (Cheby.mainstep!) */
#define TL_K_PPCG64 7  /* PPCG inner Chebyshev step */
#define TL_K_JACOBI64 8 /* Jacobi sweep (the sweeps that do not also form the residual) */
#define TL_K_RESID64 9 /* residual! r = u0 - A u (also the FP64 refresh of TL_FP32_IR) */
#define TL_K_CGF32 10  /* FP32 inner CG iteration as one kernel (TL_FP32_KERNELS_FUSED) */
#define TL_K_PPCG2_64 11 /* two PPCG inner steps in one sweep (TL_PPCG_PAIRS_AUTO) */
#define TL_K_CHEBY2_64 12 /* two Chebyshev iterations in one sweep (TL_CHEBY_PAIRS_AUTO) */
#define TL_K_CGP_G32 13 /* TL_FP32_KERNELS_PAIRS: a pair's first sweep (the six sums) */
#define TL_K_CGP_U32 14 /* TL_FP32_KERNELS_PAIRS: a pair's second sweep (r, p, x updates) */
#define TL_K_COUNT 15
/* Bracket every period-th launch of each kernel in kernel_mask (bit =
 * 1 << TL_K_*; period >= 1) with HIP events on the context's stream (an event
 * pair costs a few microseconds of GPU time, so a period of about 16 keeps the
 * measurement out of the timing); tl_profile_read returns one kernel's
 * sampled per-launch durations (ms, launch order; launches made after a
 * solve's stop flag exit at once and are near-empty).  No counterpart in the
 * reference (it has no timers). */
int tl_profile_begin(tl_ctx *ctx, unsigned kernel_mask, int period);
int tl_profile_read(tl_ctx *ctx, int kernel_id, double *durations_ms, int max, int *count);
int tl_profile_end(tl_ctx *ctx);
/* For counter collection and timing on the current fields: kernel_id of a
 * CG kernel runs `reps` iterations of that CG (FP64 or FP32 inner; both of
 * its kernels per iteration, after three warm-up iterations, with a stop test
 * that never fires), TL_K_CG_P64 `reps` launches of that kernel; -1 runs a
 * copy of one field's (nx+2h)*(ny+2h) cells (8 B read + 8 B written per cell,
 * one cell per lane), -2 a 16-byte-per-lane copy of one field's whole
 * allocation; TL_K_CHEBY64, TL_K_PPCG64, TL_K_JACOBI64 and TL_K_RESID64 run
 * `reps` launches of that kernel on the CG-initialised fields with fixed
 * coefficients (values not meaningful: for timing and counters only);
 * TL_K_CGP_G32 and TL_K_CGP_U32 run `reps` pairs of FP32 inner iterations
 * (both sweeps of each pair, after two warm-up pairs) */
int tl_kernel_repeat(tl_ctx *ctx, int kernel_id, int reps);

/* --------------------------------------------------- learned warm start */
/* BF16 U-Net (3 levels, 16/32/64 channels, 3x3 convs, FP32 accumulation)
 * predicting the timestep's change of u from (u_prev, density, energy, Kx, Ky)
 * and the residual r0 = u_prev - A u_prev of the previous-step guess.
 * No counterpart in the reference. */
typedef struct {
    int steps;            /* optimiser steps */
    int batch;            /* crops per step */
    int crop;             /* crop edge in cells (multiple of 8) */
    float lr;             /* Adam learning rate */
    int seed;
    int reserved[8];      /* reserved[0]: loss on the error e = pred - target: 0 = residual ||A e||^2,
                             1 = squared error ||e||^2, 2 = energy norm e^T A e (what CG minimises) */
} tl_train_params;

/* number of FP32 parameters of the network (flat layout per layer:
 * W[tap][cout][cin] with tap = 3*(dy+1)+(dx+1), then bias[cout]) */
int tl_unet_param_count(void);
/* load / read the network's FP32 master weights (host, tl_unet_param_count entries) */
int tl_unet_set_weights(tl_ctx *ctx, const float *w);
int tl_unet_get_weights(tl_ctx *ctx, float *w);
/* in-device dataset recorded from this context's own solves: after every
 * tl_solve, crops_per_step random 128x128 crops of (network input, target
 * (u_new - u_prev)/s_u), up to `capacity` samples; capacity 0 stops recording */
int tl_unet_record(tl_ctx *ctx, int capacity, int crops_per_step);
int tl_unet_dataset_size(tl_ctx *ctx);
/* Adam steps on minibatches of the dataset; loss_out (may be NULL) gets one
 * mean squared error per step */
int tl_unet_train(tl_ctx *ctx, const tl_train_params *tp, float *loss_out);
/* the warm-start prediction of u for the current fields into TL_F_MI */
int tl_unet_predict(tl_ctx *ctx, double rx, double ry, int coefficient);
/* diagnostics: network output (and, with target and grads, the MSE loss and
 * the gradient w.r.t. the flat parameters) for host inputs x[B][H][W][6] */
int tl_unet_eval(tl_ctx *ctx, const float *x, const float *target, int B, int H, int W, float *out, float *loss,
                 float *grads);

#ifdef __cplusplus
}
#endif
#endif
