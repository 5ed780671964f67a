/*
 * tealeaf_oracle.c -- FP64 CPU restatement of TeaLeaf.jl's linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for libtealeaf_hip.  Nothing in
 * the product links or calls this file.
 *
 * Default ("intended") mode follows the reference algorithms with the
 * mechanical porting defects corrected; each OR_LIT_* bit reproduces one of
 * them as written.  Citations are to files of Laura7089/TeaLeaf.jl.
 *
 * Reductions are deterministic for any thread count: each row of the grid is
 * summed sequentially, and the row sums are combined by a fixed pairwise tree.
 */
#include "tealeaf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ERROR_START 1e10      /* kernels.jl:7 */
#define ERROR_SWITCH_MAX 1.0  /* kernels.jl:8 */
#define CGEIGENITERS 20       /* Cheby.jl:7 */

struct or_chunk {
    or_settings s;
    int X, Y, hd;
    double *f[OR_NFIELDS];
    double *vx, *vy, *cx, *cy;          /* vertices and cell centres, chunk.jl:40-44 */
    double *hist[4];                    /* cgα, cgβ, chebyα, chebyβ (1-based), chunk.jl:56-59 */
    double theta, eigmin, eigmax;       /* chunk.jl:52-54 */
    double *rowsum;
    unsigned toex;                      /* settings.toexchange, settings.jl:46-47 */
    int maxalloc;                       /* length of the histories */
};

#define AT(c, k, j) ((size_t)(j) * (size_t)(c)->X + (size_t)(k))

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* fixed pairwise tree over n partial sums */
static double pairwise(const double *v, int n)
{
    if (n <= 0) return 0.0;
    if (n == 1) return v[0];
    if (n == 2) return v[0] + v[1];
    int h = n / 2;
    return pairwise(v, h) + pairwise(v + h, n - h);
}

/* ---------------------------------------------------------------- chunk */

or_chunk *or_create(const or_settings *s)
{
    or_chunk *c = (or_chunk *)calloc(1, sizeof(or_chunk));
    if (!c) return NULL;
    c->s = *s;
    c->hd = s->halodepth;
    c->X = s->xcells + 2 * s->halodepth;   /* chunk.jl:69 */
    c->Y = s->ycells + 2 * s->halodepth;   /* chunk.jl:70 */
    size_t n = (size_t)c->X * (size_t)c->Y;
    for (int i = 0; i < OR_NFIELDS; i++) {
        c->f[i] = (double *)calloc(n, sizeof(double));
        if (!c->f[i]) { or_destroy(c); return NULL; }
    }
    c->vx = (double *)calloc((size_t)c->X + 1, sizeof(double));
    c->vy = (double *)calloc((size_t)c->Y + 1, sizeof(double));
    c->cx = (double *)calloc((size_t)c->X, sizeof(double));
    c->cy = (double *)calloc((size_t)c->Y, sizeof(double));
    c->rowsum = (double *)calloc((size_t)c->Y, sizeof(double));
    for (int i = 0; i < 4; i++) c->hist[i] = (double *)calloc((size_t)s->maxiters + 2, sizeof(double));
    c->maxalloc = s->maxiters;
    /* vertexx = xmin + dx*((1:x+1) - hd - 1), chunk.jl:76-77 */
    for (int i = 0; i <= c->X; i++) c->vx[i] = s->xmin + s->dx * (double)(i - c->hd);
    for (int i = 0; i <= c->Y; i++) c->vy[i] = s->ymin + s->dy * (double)(i - c->hd);
    /* cell centres, chunk.jl:43-44 (celly from vertexy: D5 corrected) */
    for (int i = 0; i < c->X; i++) c->cx[i] = 0.5 * (c->vx[i] + c->vx[i + 1]);
    for (int i = 0; i < c->Y; i++) c->cy[i] = 0.5 * (c->vy[i] + c->vy[i + 1]);
    return c;
}

void or_destroy(or_chunk *c)
{
    if (!c) return;
    for (int i = 0; i < OR_NFIELDS; i++) free(c->f[i]);
    for (int i = 0; i < 4; i++) free(c->hist[i]);
    free(c->vx); free(c->vy); free(c->cx); free(c->cy); free(c->rowsum);
    free(c);
}

int or_dims(const or_chunk *c, int *X, int *Y) { *X = c->X; *Y = c->Y; return c->hd; }

/* lower the iteration cap (the CG history was sized by the settings' maxiters) */
int or_set_maxiters(or_chunk *c, int n)
{
    if (n < 1) return 1;
    if (n > c->maxalloc) return 2;
    c->s.maxiters = n;
    return 0;
}
double *or_field(or_chunk *c, int id) { return (id >= 0 && id < OR_NFIELDS) ? c->f[id] : NULL; }
double *or_history(or_chunk *c, int which) { return (which >= 0 && which < 4) ? c->hist[which] : NULL; }

/* ----------------------------------------------------------------- halo */

static double *exch_field(or_chunk *c, unsigned bit)
{
    switch (bit) {
    case OR_X_DENSITY: return c->f[OR_F_DENSITY];
    case OR_X_P: return c->f[OR_F_P];
    case OR_X_ENERGY0: return c->f[OR_F_ENERGY0];
    case OR_X_ENERGY: return c->f[OR_F_ENERGY];
    case OR_X_U: return c->f[OR_F_U];
    case OR_X_SD: return c->f[OR_F_SD];
    }
    return NULL;
}

/* updateface!, kernels.jl:191-210.  Intended (D2 corrected): reflective copy
 * g[hd+1-kk] = g[hd+kk] and g[X-hd+kk] = g[X-hd+1-kk] (1-based), over the
 * interior span only, so corners are untouched. */
static void update_face(or_chunk *c, int depth, double *g)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    const int lit = (c->s.literal & OR_LIT_D2) != 0;
    for (int kk = 1; kk <= depth; kk++) {
        int dl = lit ? hd - kk - 1 : hd - kk;        /* 0-based destination, left */
        int sl = lit ? hd + kk - 1 : hd + kk - 1;    /* 0-based source, left */
        int dr = X - hd + kk - 1;                    /* 0-based destination, right */
        int sr = lit ? X - hd - 2 - kk : X - hd - kk;
        if (dl < 0 || sr < 0) continue;
        for (int j = hd; j < Y - hd; j++) {
            g[AT(c, dl, j)] = g[AT(c, sl, j)];
            g[AT(c, dr, j)] = g[AT(c, sr, j)];
        }
    }
    for (int jj = 1; jj <= depth; jj++) {
        int db = lit ? hd - jj - 1 : hd - jj;
        int sb = hd + jj - 1;
        int dt = Y - hd + jj - 1;
        int st = lit ? Y - hd - 2 - jj : Y - hd - jj;
        if (db < 0 || st < 0) continue;
        for (int k = hd; k < X - hd; k++) {
            g[AT(c, k, dt)] = g[AT(c, k, st)];
            g[AT(c, k, db)] = g[AT(c, k, sb)];
        }
    }
}

/* haloupdate!(chunk, set, depth, toex, reset), kernels.jl:146-159 */
static void halo_explicit(or_chunk *c, unsigned list, int depth, int reset)
{
    if (reset) c->toex = 0;
    c->toex |= list;
    for (unsigned bit = 1; bit <= OR_X_SD; bit <<= 1)
        if (list & bit) update_face(c, depth, exch_field(c, bit));
}

/* haloupdate!(chunk, set, depth) with set.toexchange unchanged */
static void halo_flagged(or_chunk *c, int depth)
{
    for (unsigned bit = 1; bit <= OR_X_SD; bit <<= 1)
        if (c->toex & bit) update_face(c, depth, exch_field(c, bit));
}

void or_halo_update(or_chunk *c, unsigned mask, int depth) { halo_explicit(c, mask, depth, 1); }

/* ------------------------------------------------------------- operator */

/* smvp, kernels.jl:178-184.  Intended: the full 5-point form of the dead
 * lines 182-183, association as written.  OR_LIT_D1: diagonal term only. */
static inline double smvp(const or_chunk *c, const double *a, size_t i)
{
    const size_t X = (size_t)c->X;
    const double *kx = c->f[OR_F_KX], *ky = c->f[OR_F_KY];
    double consum = 1.0 + kx[i + 1];
    consum = consum + kx[i];
    consum = consum + ky[i + X];
    consum = consum + ky[i];
    if (c->s.literal & OR_LIT_D1) return consum * a[i];
    return (consum * a[i] - (kx[i + 1] * a[i + 1] + kx[i] * a[i - 1])) -
           (ky[i + X] * a[i + X] + ky[i] * a[i - X]);
}

void or_smvp_field(or_chunk *c, int src, int dst)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    const double *a = c->f[src];
    double *o = c->f[dst];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) o[AT(c, k, j)] = smvp(c, a, AT(c, k, j));
}

/* Σ over the interior of f(i), per-row sequential, pairwise over rows */
#define INTERIOR_SUM(c, expr, out)                                              \
    do {                                                                        \
        const int X_ = (c)->X, Y_ = (c)->Y, hd_ = (c)->hd;                      \
        _Pragma("omp parallel for schedule(static)")                            \
        for (int j = hd_; j < Y_ - hd_; j++) {                                  \
            double t_ = 0.0;                                                    \
            for (int k = hd_; k < X_ - hd_; k++) {                              \
                const size_t i = AT(c, k, j);                                   \
                (void)i;                                                        \
                t_ += (expr);                                                   \
            }                                                                   \
            (c)->rowsum[j - hd_] = t_;                                          \
        }                                                                       \
        (out) = pairwise((c)->rowsum, Y_ - 2 * hd_);                            \
    } while (0)

/* ---------------------------------------------------------- initialise */

/* setchunkstate!, chunk.jl:122-151 */
static int set_chunk_state(or_chunk *c)
{
    const int X = c->X, Y = c->Y;
    const size_t n = (size_t)X * (size_t)Y;
    double *e0 = c->f[OR_F_ENERGY0], *rho = c->f[OR_F_DENSITY];
    const or_settings *s = &c->s;
    if (s->nstates < 1) return 1;
    for (size_t i = 0; i < n; i++) { e0[i] = s->states[0].energy; rho[i] = s->states[0].density; }
    int bad = 0;
    for (int si = 0; si < s->nstates; si++) {
        const or_state *st = &s->states[si];
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (int j = 0; j < Y; j++) {
            for (int k = 0; k < X; k++) {
                int apply = 0;
                if (st->geometry == OR_RECT) {
                    apply = c->vx[k + 1] >= st->xmin && c->vx[k] < st->xmax &&
                            c->vy[j + 1] >= st->ymin && c->vy[j] < st->ymax;
                } else if (st->geometry == OR_CIRC) {
                    double ddx = c->cx[k] - st->xmin, ddy = c->cy[j] - st->ymin;
                    apply = ddx * ddx + ddy * ddy <= st->radius * st->radius;
                } else {
                    apply = c->vx[k] == st->xmin && c->vy[j] == st->ymin;
                }
                if (!apply) continue;
                size_t idx;
                if (s->literal & OR_LIT_D4) {
                    /* ch.energy0[j, k]: x-index j, y-index k */
                    if (j >= X || k >= Y) { bad = 1; continue; } /* BoundsError in the reference */
                    idx = (size_t)k * (size_t)X + (size_t)j;
                } else {
                    idx = AT(c, k, j);
                }
                e0[idx] = st->energy;
                rho[idx] = st->density;
            }
        }
    }
    if (bad) return 2;
    /* u = energy0*density on halo(chunk, 1), chunk.jl:149-150 */
    double *u = c->f[OR_F_U];
    for (int j = 1; j < Y - 1; j++)
        for (int k = 1; k < X - 1; k++) u[AT(c, k, j)] = e0[AT(c, k, j)] * rho[AT(c, k, j)];
    return 0;
}

/* initialiseapp!, TeaLeaf.jl:35-44 */
int or_initialise(or_chunk *c)
{
    int rc = set_chunk_state(c);
    if (rc) return rc;
    halo_explicit(c, OR_X_DENSITY | OR_X_ENERGY0 | OR_X_ENERGY, 1, 1);
    memcpy(c->f[OR_F_ENERGY], c->f[OR_F_ENERGY0], sizeof(double) * (size_t)c->X * (size_t)c->Y);
    return 0;
}

/* ------------------------------------------------------------------ CG */

/* coefficient fields, CG.jl:61-68 (and Jacobi.jl:47-55, same formula) */
static void calc_coefficients(or_chunk *c, const double *w, double rx, double ry)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *kx = c->f[OR_F_KX], *ky = c->f[OR_F_KY];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - 1; j++) {
        for (int k = hd; k < X - 1; k++) {
            const size_t i = AT(c, k, j);
            kx[i] = rx * (w[i - 1] + w[i]) / (2.0 * w[i - 1] * w[i]);
            ky[i] = ry * (w[i - X] + w[i]) / (2.0 * w[i - X] * w[i]);
        }
    }
}

/* CG.init!, CG.jl:47-79 */
double or_cg_init(or_chunk *c, double rx, double ry, int *status)
{
    const int X = c->X, Y = c->Y;
    const size_t n = (size_t)X * (size_t)Y;
    int coef = c->s.coefficient;
    if (coef != 1 && coef != 2) { *status = 1; return 0.0; }
    *status = 0;
    double *u = c->f[OR_F_U], *e = c->f[OR_F_ENERGY], *rho = c->f[OR_F_DENSITY];
    double *p = c->f[OR_F_P], *r = c->f[OR_F_R], *w = c->f[OR_F_W];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) { u[i] = e[i] * rho[i]; p[i] = 0.0; r[i] = 0.0; }
    for (int j = 1; j < Y - 1; j++)
        for (int k = 1; k < X - 1; k++) {
            const size_t i = AT(c, k, j);
            w[i] = coef == 1 ? rho[i] : 1.0 / rho[i];
        }
    calc_coefficients(c, w, rx, ry);
    double temp;
    {
        const int hd = c->hd;
#pragma omp parallel for schedule(static)
        for (int j = hd; j < Y - hd; j++) {
            double t = 0.0;
            for (int k = hd; k < X - hd; k++) {
                const size_t i = AT(c, k, j);
                w[i] = smvp(c, u, i);
                r[i] = u[i] - w[i];
                p[i] = r[i];
                t += r[i] * r[i];
            }
            c->rowsum[j - hd] = t;
        }
        temp = pairwise(c->rowsum, Y - 2 * hd);
    }
    return temp;
}

/* CG.w!, CG.jl:82-90 */
double or_cg_w(or_chunk *c)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *w = c->f[OR_F_W];
    const double *p = c->f[OR_F_P];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++) {
        double t = 0.0;
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            w[i] = smvp(c, p, i);
            t += w[i] * p[i];
        }
        c->rowsum[j - hd] = t;
    }
    return pairwise(c->rowsum, Y - 2 * hd);
}

/* CG.ur!, CG.jl:93-98; the sum runs over the whole r, whose halo is zero */
double or_cg_ur(or_chunk *c, double alpha)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *u = c->f[OR_F_U], *r = c->f[OR_F_R];
    const double *p = c->f[OR_F_P], *w = c->f[OR_F_W];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++) {
        double t = 0.0;
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            u[i] += alpha * p[i];
            r[i] -= alpha * w[i];
            t += r[i] * r[i];
        }
        c->rowsum[j - hd] = t;
    }
    return pairwise(c->rowsum, Y - 2 * hd);
}

/* CG.p!, CG.jl:101-104 */
void or_cg_p(or_chunk *c, double beta)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *p = c->f[OR_F_P];
    const double *r = c->f[OR_F_R];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            p[i] = beta * p[i] + r[i];
        }
}

/* copyu!, kernels.jl:217-220 */
static void copyu(or_chunk *c)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) c->f[OR_F_U0][AT(c, k, j)] = c->f[OR_F_U][AT(c, k, j)];
}

/* CG.mainstep!, CG.jl:32-44 */
static double cg_mainstep(or_chunk *c, int tt, double rro)
{
    double pw = or_cg_w(c);
    double alpha = rro / pw;
    c->hist[0][tt] = alpha;
    double rrn = or_cg_ur(c, alpha);
    double beta = rrn / rro;
    c->hist[1][tt] = beta;
    or_cg_p(c, beta);
    return rrn;
}

/* CG-init shared by every solver: init!, halo(u,p), copyu! (CG.jl:9-12; D18) */
static double cg_init_full(or_chunk *c, double rx, double ry, int *status)
{
    double rro = or_cg_init(c, rx, ry, status);
    if (*status) return 0.0;
    halo_explicit(c, OR_X_U | OR_X_P, 1, 1);
    copyu(c);
    return rro;
}

static double bnorm(or_chunk *c)
{
    double bb;
    const double *u0 = c->f[OR_F_U0];
    INTERIOR_SUM(c, u0[i] * u0[i], bb);
    return sqrt(bb);
}

static int cg_stop(const or_chunk *c, double error, double bn)
{
    if (c->s.stop_rule == OR_STOP_RELATIVE) return sqrt(fabs(error)) <= c->s.rel_tol * bn;
    return sqrt(fabs(error)) < c->s.eps; /* CG.jl:24 */
}

/* CG.solve!, CG.jl:7-29 */
static int cg_solve(or_chunk *c, double rx, double ry, or_stats *st)
{
    int status;
    double rro = cg_init_full(c, rx, ry, &status);
    if (status) return status;
    double bn = bnorm(c);
    double error = ERROR_START;
    int iters = 0;
    /* breakdown guard: an exactly zero initial residual is already the solution */
    if (rro != 0.0) {
        for (int tt = 1; tt <= c->s.maxiters; tt++) {
            iters = tt;
            rro = error = cg_mainstep(c, tt, rro);
            halo_flagged(c, 1);
            if (cg_stop(c, error, bn)) break;
        }
    } else {
        error = 0.0;
    }
    st->iterations = iters;
    st->cg_iters = iters;
    st->error = error;
    return 0;
}

/* ------------------------------------------------------ eigen / coefficients */

/* Eigenvalues of the symmetric tridiagonal (d, e): implicit QL with Wilkinson
 * shifts (the tqli routine of Numerical Recipes 11.3, which kernels.jl:55-112
 * cites).  e[i] couples d[i] and d[i+1]; e[n-1] is ignored. */
int or_tqli(double *d, double *e, int n)
{
    if (n <= 0) return 1;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) return 2;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + copysign(r, g));
                double s = 1.0, cc = 1.0, p = 0.0;
                int i, underflow = 0;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i], b = cc * e[i];
                    r = hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        underflow = 1;
                        break;
                    }
                    s = f / r;
                    cc = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * cc * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = cc * r - b;
                }
                if (underflow) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return 0;
}

/* eigenvalues!, kernels.jl:19-51 with D9-D11 corrected: the Lanczos
 * tridiagonal of n CG steps, then extrema scaled by 0.95 / 1.05.
 * cga/cgb are 1-based (index 1..n). */
int or_eigenvalues(const double *cga, const double *cgb, int n, double *eigmin, double *eigmax)
{
    if (n < 1) return 3;
    double *d = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)calloc((size_t)n, sizeof(double));
    for (int ii = 1; ii <= n; ii++) {
        d[ii - 1] = 1.0 / cga[ii];
        if (ii > 1) d[ii - 1] += cgb[ii - 1] / cga[ii - 1];
        if (ii < n) e[ii - 1] = sqrt(cgb[ii]) / cga[ii];
    }
    int rc = or_tqli(d, e, n);
    if (rc == 0) {
        double mn = d[0], mx = d[0];
        for (int i = 1; i < n; i++) { if (d[i] < mn) mn = d[i]; if (d[i] > mx) mx = d[i]; }
        if (mn < 0 || mx < 0) rc = 4; /* kernels.jl:41-43 */
        *eigmin = mn * 0.95;
        *eigmax = mx * 1.05;
    }
    free(d);
    free(e);
    return rc;
}

/* Cheby.coef!, Cheby.jl:121-135; alpha/beta are 1-based (1..n) */
void or_cheby_coef(double eigmin, double eigmax, int n, double *alpha, double *beta, double *theta)
{
    double th = (eigmax + eigmin) / 2.0;
    double delta = (eigmax - eigmin) / 2.0;
    double sigma = th / delta;
    double rho_old = 1.0 / sigma;
    for (int ii = 1; ii <= n; ii++) {
        double rho_new = 1.0 / (2.0 * sigma - rho_old);
        alpha[ii] = rho_new * rho_old;
        beta[ii] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
    *theta = th;
}

/* Cheby.calciter, Cheby.jl:109-118 (Julia round: nearest, ties to even) */
int or_calciter(double eigmin, double eigmax, double error, double bb)
{
    double connum = eigmax / eigmin;
    double ita = DBL_EPSILON * bb / (4.0 * error);
    double gamma = (sqrt(connum) - 1.0) / (sqrt(connum) + 1.0);
    double v = (log(ita) / log(10.0)) / (2.0 * (log(gamma) / log(10.0)));
    if (!isfinite(v)) return -1;
    return (int)rint(v);
}

/* ---------------------------------------------------------------- Cheby */

/* Cheby.init!, Cheby.jl:64-81 (cgiters = tt-1: D9) */
static int cheby_init(or_chunk *c, int cgiters, double *bb)
{
    int rc = or_eigenvalues(c->hist[0], c->hist[1], cgiters, &c->eigmin, &c->eigmax);
    if (rc) return rc;
    or_cheby_coef(c->eigmin, c->eigmax, c->s.maxiters - cgiters, c->hist[2], c->hist[3], &c->theta);
    const double *u0 = c->f[OR_F_U0];
    INTERIOR_SUM(c, u0[i] * u0[i], *bb);
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *u = c->f[OR_F_U], *w = c->f[OR_F_W], *r = c->f[OR_F_R], *p = c->f[OR_F_P];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            w[i] = smvp(c, u, i);
            r[i] = u0[i] - w[i];
            p[i] = r[i] / c->theta;
        }
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) u[AT(c, k, j)] += p[AT(c, k, j)];
    halo_explicit(c, OR_X_U, 1, 1);
    return 0;
}

/* Cheby.mainstep!, Cheby.jl:84-106 */
static double cheby_mainstep(or_chunk *c, int chebyiters, int calc2norm, double error)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *u = c->f[OR_F_U], *w = c->f[OR_F_W], *r = c->f[OR_F_R], *p = c->f[OR_F_P];
    const double *u0 = c->f[OR_F_U0];
    const double a = c->hist[2][chebyiters + 1], b = c->hist[3][chebyiters + 1];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            w[i] = smvp(c, u, i);
            r[i] = u0[i] - w[i];
            p[i] = a * p[i] + b * r[i];
        }
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) u[AT(c, k, j)] += p[AT(c, k, j)];
    if (calc2norm) INTERIOR_SUM(c, r[i] * r[i], error);
    return error;
}

static int switch_test(const or_chunk *c, int tt, double error)
{
    /* Cheby.jl:25-29, PPCG.jl:25-30 */
    return c->s.errorswitch ? (error < c->s.epslim && tt > CGEIGENITERS)
                            : (tt > c->s.presteps && error < ERROR_SWITCH_MAX);
}

/* Cheby.solve!, Cheby.jl:10-61 with D8, D9, D11, D18 corrected */
static int cheby_solve(or_chunk *c, double rx, double ry, or_stats *st)
{
    int status;
    double error = ERROR_START;
    int estiter = 0, chebyiters = 0, iters = 0, cgiters = 0;
    double rro = cg_init_full(c, rx, ry, &status);
    if (status) return status;
    if (rro == 0.0) { st->error = 0.0; return 0; }
    for (int tt = 1; tt <= c->s.maxiters; tt++) {
        iters = tt;
        int sw = chebyiters != 0 || switch_test(c, tt, error);
        if (!sw) {
            rro = error = cg_mainstep(c, tt, rro);
            cgiters = tt;
        } else {
            chebyiters++;
            if (chebyiters == 1) {
                double bb;
                int rc = cheby_init(c, tt - 1, &bb);
                if (rc) return rc;
                error = cheby_mainstep(c, chebyiters, 1, error);
                estiter = or_calciter(c->eigmin, c->eigmax, error, bb);
            } else {
                int calc2norm = chebyiters >= estiter && (tt + 1) % 10 == 0;
                error = cheby_mainstep(c, chebyiters, calc2norm, error);
            }
        }
        halo_flagged(c, 1);
        if (fabs(error) < c->s.eps) break;
    }
    st->iterations = iters;
    st->cg_iters = cgiters;
    st->cheby_iters = chebyiters;
    st->est_iters = estiter;
    st->error = error;
    st->eigmin = c->eigmin;
    st->eigmax = c->eigmax;
    return 0;
}

/* ----------------------------------------------------------------- PPCG */

/* residual!, kernels.jl:227-232 */
void or_residual(or_chunk *c)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *r = c->f[OR_F_R];
    const double *u0 = c->f[OR_F_U0], *u = c->f[OR_F_U];
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            r[i] = u0[i] - smvp(c, u, i);
        }
}

/* PPCG.mainstep!, PPCG.jl:64-94 with D14-D16 corrected */
static double ppcg_mainstep(or_chunk *c, double rro)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double pw = or_cg_w(c);
    double alpha = rro / pw;
    (void)or_cg_ur(c, alpha);
    double *r = c->f[OR_F_R], *u = c->f[OR_F_U], *sd = c->f[OR_F_SD];
    /* PPCG.init!(chunk, hd): sd = r/θ (PPCG.jl:97-100; D15) */
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) sd[AT(c, k, j)] = r[AT(c, k, j)] / c->theta;
    c->toex = OR_X_SD;
    for (int pp = 1; pp <= c->s.ppcginnersteps; pp++) {
        halo_flagged(c, 1);
        /* two phases (D16): finish r and u for every cell, then update sd */
#pragma omp parallel for schedule(static)
        for (int j = hd; j < Y - hd; j++)
            for (int k = hd; k < X - hd; k++) {
                const size_t i = AT(c, k, j);
                r[i] -= smvp(c, sd, i);
                u[i] += sd[i];
            }
        const double a = c->hist[2][pp], b = c->hist[3][pp];
#pragma omp parallel for schedule(static)
        for (int j = hd; j < Y - hd; j++)
            for (int k = hd; k < X - hd; k++) {
                const size_t i = AT(c, k, j);
                sd[i] = a * sd[i] + b * r[i];
            }
    }
    c->toex = OR_X_P;
    double rrn;
    INTERIOR_SUM(c, r[i] * r[i], rrn);
    double beta = rrn / rro;
    or_cg_p(c, beta);
    return rrn;
}

/* PPCG.solve!, PPCG.jl:9-55 with D9, D12, D13, D18 corrected */
static int ppcg_solve(or_chunk *c, double rx, double ry, or_stats *st)
{
    int status, ppcgiters = 0, iters = 0, cgiters = 0;
    double error = ERROR_START;
    double rro = cg_init_full(c, rx, ry, &status);
    if (status) return status;
    if (rro == 0.0) { st->error = 0.0; return 0; }
    for (int tt = 1; tt <= c->s.maxiters; tt++) {
        iters = tt;
        int sw = ppcgiters != 0 || switch_test(c, tt, error);
        if (!sw) {
            rro = error = cg_mainstep(c, tt, rro);
            cgiters = tt;
        } else {
            ppcgiters++;
            if (ppcgiters == 1) {
                int rc = or_eigenvalues(c->hist[0], c->hist[1], tt - 1, &c->eigmin, &c->eigmax);
                if (rc) return rc;
                or_cheby_coef(c->eigmin, c->eigmax, c->s.ppcginnersteps, c->hist[2], c->hist[3], &c->theta);
                /* PPCG.init!(chunk, settings), PPCG.jl:58-61 */
                or_residual(c);
                halo_explicit(c, OR_X_P, 1, 1);
            }
            rro = error = ppcg_mainstep(c, rro);
        }
        halo_flagged(c, 1);
        if (fabs(error) < c->s.eps) break;
    }
    st->iterations = iters;
    st->cg_iters = cgiters;
    st->cheby_iters = ppcgiters;
    st->error = error;
    st->eigmin = c->eigmin;
    st->eigmax = c->eigmax;
    return 0;
}

/* --------------------------------------------------------------- Jacobi */

/* Jacobi.init!, Jacobi.jl:33-57 */
static int jacobi_init(or_chunk *c, double rx, double ry)
{
    int coef = c->s.coefficient;
    if (coef != 1 && coef != 2) return 1;
    const size_t n = (size_t)c->X * (size_t)c->Y;
    double *u = c->f[OR_F_U], *u0 = c->f[OR_F_U0], *w = c->f[OR_F_W];
    const double *e = c->f[OR_F_ENERGY], *rho = c->f[OR_F_DENSITY];
    for (size_t i = 0; i < n; i++) { u0[i] = e[i] * rho[i]; u[i] = u0[i]; }
    /* densc = density^p: staged in w over halo(1) as CG.init! does */
    for (int j = 1; j < c->Y - 1; j++)
        for (int k = 1; k < c->X - 1; k++) {
            const size_t i = AT(c, k, j);
            w[i] = coef == 1 ? rho[i] : 1.0 / rho[i];
        }
    calc_coefficients(c, w, rx, ry);
    copyu(c);
    c->toex = OR_X_U;
    return 0;
}

/* Jacobi.iterate!, Jacobi.jl:59-80 */
static double jacobi_iterate(or_chunk *c)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    const size_t n = (size_t)X * (size_t)Y;
    double *u = c->f[OR_F_U], *r = c->f[OR_F_R];
    const double *u0 = c->f[OR_F_U0], *kx = c->f[OR_F_KX], *ky = c->f[OR_F_KY];
    memcpy(r, u, n * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int j = hd; j < Y - hd; j++) {
        double t = 0.0;
        for (int k = hd; k < X - hd; k++) {
            const size_t i = AT(c, k, j);
            double num = u0[i] + kx[i + 1] * r[i + 1];
            num = num + kx[i] * r[i - 1];
            num = num + ky[i + X] * r[i + X];
            num = num + ky[i] * r[i - X];
            double den = 1.0 + kx[i];
            den = den + kx[i + 1];
            den = den + ky[i];
            den = den + ky[i + X];
            u[i] = num / den;
            t += fabs(u[i] - r[i]);
        }
        c->rowsum[j - hd] = t;
    }
    return pairwise(c->rowsum, Y - 2 * hd);
}

/* Jacobi.driver!, Jacobi.jl:7-31 with D17 corrected */
static int jacobi_solve(or_chunk *c, double rx, double ry, or_stats *st)
{
    if (jacobi_init(c, rx, ry)) return 1;
    double error = ERROR_START;
    int iters = 0;
    for (int tt = 1; tt <= c->s.maxiters; tt++) {
        iters = tt;
        error = jacobi_iterate(c);
        if (tt % 50 == 0) {
            halo_flagged(c, 1);
            or_residual(c);
            double rr;
            const double *r = c->f[OR_F_R];
            INTERIOR_SUM(c, r[i] * r[i], rr);
            error += rr;
        }
        halo_flagged(c, 1);
        if (fabs(error) < c->s.eps) break;
    }
    st->iterations = iters;
    st->error = error;
    return 0;
}

/* ------------------------------------------------------ step and summary */

/* finalise!, kernels.jl:239-242 */
void or_finalise(or_chunk *c)
{
    const int X = c->X, Y = c->Y, hd = c->hd;
    double *e = c->f[OR_F_ENERGY];
    const double *u = c->f[OR_F_U], *rho = c->f[OR_F_DENSITY];
    for (int j = hd; j < Y - hd; j++)
        for (int k = hd; k < X - hd; k++) e[AT(c, k, j)] = u[AT(c, k, j)] / rho[AT(c, k, j)];
}

/* fieldsummary, kernels.jl:119-133: temp = Σ vol·ρ·u; vol, mass and ie
 * (Σ vol, Σ vol·ρ, Σ vol·ρ·energy) are reported alongside */
void or_field_summary(or_chunk *c, double out[4])
{
    const double vol = c->s.dx * c->s.dy; /* chunk.jl:79 */
    const double *rho = c->f[OR_F_DENSITY], *e = c->f[OR_F_ENERGY], *u = c->f[OR_F_U];
    double v, m, ie, t;
    INTERIOR_SUM(c, vol, v);
    INTERIOR_SUM(c, vol * rho[i], m);
    INTERIOR_SUM(c, vol * rho[i] * e[i], ie);
    INTERIOR_SUM(c, vol * rho[i] * u[i], t);
    out[0] = v; out[1] = m; out[2] = ie; out[3] = t;
}

int or_solve(or_chunk *c, double rx, double ry, or_stats *st)
{
    memset(st, 0, sizeof(*st));
    int rc;
    switch (c->s.solver) {
    case OR_CG: rc = cg_solve(c, rx, ry, st); break;
    case OR_CHEBY: rc = cheby_solve(c, rx, ry, st); break;
    case OR_PPCG: rc = ppcg_solve(c, rx, ry, st); break;
    case OR_JACOBI: rc = jacobi_solve(c, rx, ry, st); break;
    default: rc = 1;
    }
    st->status = rc;
    if (rc == 0) {
        /* true relative residual ||u0 - A u|| / ||u0|| of the returned u */
        const double *u0 = c->f[OR_F_U0], *u = c->f[OR_F_U];
        double rr, bb;
        INTERIOR_SUM(c, (u0[i] - smvp(c, u, i)) * (u0[i] - smvp(c, u, i)), rr);
        INTERIOR_SUM(c, u0[i] * u0[i], bb);
        st->rel_resid = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
    }
    return rc;
}

/* one timestep of diffuse!, TeaLeaf.jl:67-81 */
int or_step(or_chunk *c, or_stats *st)
{
    const double rx = c->s.dtinit / (c->s.dx * c->s.dx); /* TeaLeaf.jl:69-70 */
    const double ry = c->s.dtinit / (c->s.dy * c->s.dy);
    halo_explicit(c, OR_X_ENERGY | OR_X_DENSITY, 1, 1);
    int rc = or_solve(c, rx, ry, st);
    if (rc) return rc;
    /* solvefinished!, kernels.jl:166-170 */
    if (c->s.checkresult) or_residual(c);
    or_finalise(c);
    halo_explicit(c, OR_X_ENERGY, 1, 0);
    return 0;
}

int or_run(or_chunk *c, int nsteps, double *summary, or_stats *stats)
{
    for (int t = 0; t < nsteps; t++) {
        int rc = or_step(c, stats ? &stats[t] : &(or_stats){0});
        if (rc) return rc;
        if (summary) or_field_summary(c, summary + 4 * t);
    }
    return 0;
}
