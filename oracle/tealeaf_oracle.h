/*
 * tealeaf_oracle.h -- CPU restatement of TeaLeaf.jl's linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation (libtealeaf_hip).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Layout follows src/chunk.jl:19-89 of the reference: every field is an
 * X*Y array of doubles, X = xcells + 2*halodepth, Y = ycells + 2*halodepth,
 * column-major with x fastest (Julia Matrix{Float64}).  Index (k, j) here is
 * 0-based: element = j*X + k; the reference's 1-based [k+1, j+1].
 *
 * The reference ships no test vectors and cannot run here (Julia), so each
 * function is pinned by mathematics (tests/test_oracle.py):
 *   or_initialise        exact covered-cell counts and initial sum vol*rho*u
 *                        of the 512^2 deck (intended and literal D3/D4)
 *   or_smvp_field        cosine eigenmode (closed-form eigenvalue), symmetry,
 *                        unit row sums under reflection
 *   or_halo_update       ghost indices, intended and literal (D2)
 *   or_residual          u0 - A u against the assembled sparse matrix
 *   or_solve/or_step     CG, Chebyshev, PPCG, Jacobi: converged u equals a
 *                        sparse direct solve; cross-solver agreement; PPCG
 *                        with 0 inner steps equals CG; ie conservation
 *   or_field_summary,    uniform-deck closed form per step
 *   or_finalise
 *   or_tqli,             scipy eigvalsh_tridiagonal; Ritz values inside the
 *   or_eigenvalues       dense spectrum
 *   or_cheby_coef,       Chebyshev-polynomial closed form; the formula
 *   or_calciter
 *   literal D1           u_{n+1} = u_n / D cell by cell
 * Not pinned (parity unpinned): the reference's as-written alpha/beta path
 * (only its limit is pinned) and the rounding of Julia's pairwise sum.
 */
#ifndef TEALEAF_ORACLE_H
#define TEALEAF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_STATES 128

/* literal-mode switches: each reproduces one defect of the reference as written */
#define OR_LIT_D1 1u /* smvp returns only the diagonal term (kernels.jl:181) */
#define OR_LIT_D2 2u /* halo indices of updateface! (kernels.jl:196-208) */
#define OR_LIT_D3 4u /* state bounds offset by dx=10/100 (settings.jl:159-162); applied by the deck parser */
#define OR_LIT_D4 8u /* state image written transposed (chunk.jl:144-145) */

enum { OR_CG = 0, OR_CHEBY = 1, OR_PPCG = 2, OR_JACOBI = 3 };
enum { OR_RECT = 0, OR_CIRC = 1, OR_POINT = 2 };
enum { OR_STOP_REFERENCE = 0, OR_STOP_RELATIVE = 1 };

/* field ids (chunk.jl:25-38) */
enum {
    OR_F_DENSITY0 = 0, OR_F_DENSITY, OR_F_ENERGY0, OR_F_ENERGY, OR_F_U, OR_F_U0,
    OR_F_P, OR_F_R, OR_F_MI, OR_F_W, OR_F_KX, OR_F_KY, OR_F_SD, OR_NFIELDS
};
/* exchange mask bits (chunk.jl:12 EXCHANGE_FIELDS) */
#define OR_X_DENSITY (1u << 0)
#define OR_X_P (1u << 1)
#define OR_X_ENERGY0 (1u << 2)
#define OR_X_ENERGY (1u << 3)
#define OR_X_U (1u << 4)
#define OR_X_SD (1u << 5)

typedef struct {
    double density, energy, xmin, ymin, xmax, ymax, radius;
    int geometry;
} or_state;

typedef struct {
    int xcells, ycells, halodepth;
    double xmin, ymin, xmax, ymax, dx, dy;
    int endstep, presteps, maxiters, coefficient, ppcginnersteps, summaryfrequency;
    int errorswitch, checkresult;
    double eps, dtinit, epslim;
    int solver;
    int nstates;
    or_state states[OR_MAX_STATES];
    unsigned literal;
    int stop_rule;  /* OR_STOP_REFERENCE: each solver's own rule; OR_STOP_RELATIVE: ||r||/||u0|| <= rel_tol */
    double rel_tol;
} or_settings;

typedef struct {
    int iterations;   /* outer iterations run (tt at exit) */
    int cg_iters;     /* CG iterations before a Cheby/PPCG switch */
    int cheby_iters;  /* Chebyshev (or PPCG outer) iterations after the switch */
    int est_iters;    /* Chebyshev iteration estimate (calciter) */
    int status;       /* 0 ok, nonzero error code */
    double error;     /* final error as the solver defines it */
    double eigmin, eigmax;
    double rel_resid; /* ||u0 - A u|| / ||u0|| after the solve */
} or_stats;

typedef struct or_chunk or_chunk;

or_chunk *or_create(const or_settings *s);
void or_destroy(or_chunk *c);
int or_dims(const or_chunk *c, int *X, int *Y);
int or_set_maxiters(or_chunk *c, int n);
double *or_field(or_chunk *c, int id);
double *or_history(or_chunk *c, int which); /* 0 cgalpha 1 cgbeta 2 chebyalpha 3 chebybeta */
void or_set_threads(int n);

/* reference-level operations */
int or_initialise(or_chunk *c);                          /* initialiseapp!, TeaLeaf.jl:35-44 */
void or_halo_update(or_chunk *c, unsigned mask, int depth); /* haloupdate!, kernels.jl:146-159 */
double or_cg_init(or_chunk *c, double rx, double ry, int *status); /* CG.init!, CG.jl:47-79 */
double or_cg_w(or_chunk *c);                             /* CG.w!, CG.jl:82-90 */
double or_cg_ur(or_chunk *c, double alpha);              /* CG.ur!, CG.jl:93-98 */
void or_cg_p(or_chunk *c, double beta);                  /* CG.p!, CG.jl:101-104 */
void or_smvp_field(or_chunk *c, int src, int dst);       /* dst = A*src over the interior, kernels.jl:178-184 */
void or_residual(or_chunk *c);                           /* residual!, kernels.jl:227-232 */
void or_finalise(or_chunk *c);                           /* finalise!, kernels.jl:239-242 */
void or_field_summary(or_chunk *c, double out[4]);       /* vol, mass, ie, temp; kernels.jl:119-133 */
int or_solve(or_chunk *c, double rx, double ry, or_stats *st); /* set.solver.solve!, TeaLeaf.jl:74 */
int or_step(or_chunk *c, or_stats *st);                  /* one iteration of diffuse!, TeaLeaf.jl:67-81 */
int or_run(or_chunk *c, int nsteps, double *summary, or_stats *stats);

/* host-side scalar routines of the Chebyshev/PPCG solvers */
int or_tqli(double *d, double *e, int n);                       /* eigenvalues of a symmetric tridiagonal */
int or_eigenvalues(const double *cga, const double *cgb, int n, double *eigmin, double *eigmax);
void or_cheby_coef(double eigmin, double eigmax, int n, double *alpha, double *beta, double *theta);
int or_calciter(double eigmin, double eigmax, double error, double bb);

#ifdef __cplusplus
}
#endif
#endif
