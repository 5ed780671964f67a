// tl_api.hip -- C ABI and host-side solver drivers of libtealeaf_hip.
//
// The drivers follow the solver modules of Laura7089/TeaLeaf.jl
// (src/solvers/{CG,Cheby,PPCG,Jacobi}.jl) in their intended form.  The host
// only enqueues kernels: every scalar recurrence (alpha, beta, stop tests)
// runs on the device (tl_kernels.hip), and the host polls a snapshot of the
// device control block once per batch of iterations, one batch ahead, so
// the GPU never idles waiting for the host inside a solve.
#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tl_internal.hpp"

using namespace tl;

namespace tl {
int unet_predict_guess(tl_ctx *c, double rx, double ry, int coefficient, double *guess_dev);
void unet_destroy(tl_ctx *c);
int unet_record_step(tl_ctx *c, double rx, double ry, int coefficient);
}  // namespace tl

#define HIPCHK(c, expr)                                                                \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (c)->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return TL_E_HIP;                                                           \
        }                                                                              \
    } while (0)

#define RETIF(expr)                  \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != TL_OK) return rc_; \
    } while (0)

static thread_local std::string g_create_err;

extern "C" void tl_default_params(tl_solve_params *p)
{
    std::memset(p, 0, sizeof(*p));
    // Settings defaults, settings.jl:41-59
    p->solver = TL_CG;
    p->coefficient = 1;
    p->max_iters = 10000;
    p->presteps = 30;
    p->ppcg_inner_steps = 10;
    p->error_switch = 0;
    p->eps = 1e-15;
    p->epslim = 1e-5;
    p->precision = TL_FP64;
    p->stop_rule = TL_STOP_REFERENCE;
    p->rel_tol = 1e-10;
    p->warm_start = TL_WARM_NONE;
    p->check_result = 1;
    p->ir_delta = 0.1;
}

// ------------------------------------------------------------- lifecycle

static int alloc_d(tl_ctx *c, double **p, size_t n)
{
    if (hipMalloc((void **)p, n * sizeof(double)) != hipSuccess) {
        c->err = "device allocation failed";
        *p = nullptr;
        return TL_E_NOMEM;
    }
    if (hipMemset(*p, 0, n * sizeof(double)) != hipSuccess) return TL_E_HIP;
    return TL_OK;
}

static int alloc_f(tl_ctx *c, float **p, size_t n)
{
    if (hipMalloc((void **)p, n * sizeof(float)) != hipSuccess) {
        c->err = "device allocation failed";
        *p = nullptr;
        return TL_E_NOMEM;
    }
    if (hipMemset(*p, 0, n * sizeof(float)) != hipSuccess) return TL_E_HIP;
    return TL_OK;
}

static hipEvent_t *events(tl_ctx *c);

extern "C" int tl_create(tl_ctx **out, int nx, int ny, int halo_depth, double xmin, double ymin, double dx,
                         double dy, int device)
{
    *out = nullptr;
    // halo_depth >= 2: CG.init! (CG.jl:59-68) builds the face coefficients
    // from w on halo(chunk, 1), every cell but the outermost ring, and reads
    // w one cell outside the interior; with a depth-1 halo that cell is the
    // unset outermost ring and the reference divides by zero
    if (nx < 2 || ny < 2 || halo_depth < 2 || halo_depth > 16 || !(dx > 0) || !(dy > 0)) {
        g_create_err = "tl_create: invalid grid (nx, ny >= 2, 2 <= halo_depth <= 16, dx, dy > 0)";
        return TL_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_err = "tl_create: no HIP device";
        return TL_E_HIP;
    }
    if (device < 0 || device >= ndev) {
        g_create_err = "tl_create: invalid device";
        return TL_E_INVALID;
    }
    tl_ctx *c = new (std::nothrow) tl_ctx();
    if (!c) return TL_E_NOMEM;
    c->device = device;
    c->xmin = xmin;
    c->ymin = ymin;
    c->dx = dx;
    c->dy = dy;
    c->g = make_geo(nx, ny, halo_depth);
    int rc = TL_OK;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        rc = TL_E_HIP;
    for (int i = 0; i < TL_NFIELDS && rc == TL_OK; ++i) rc = alloc_d(c, &c->f[i], c->g.elems);
    if (rc == TL_OK) rc = alloc_d(c, &c->u_alt, c->g.elems);
    if (rc == TL_OK) rc = alloc_d(c, &c->sd_alt, c->g.elems);
    if (rc == TL_OK) rc = alloc_d(c, &c->p_alt, c->g.elems);
    c->partials_len = (size_t)7 * c->g.nbmax;  // pw | pr0 | pr1 | 4 x misc
    if (rc == TL_OK) rc = alloc_d(c, &c->partials, c->partials_len);
    if (rc == TL_OK && hipMalloc((void **)&c->ctrl, sizeof(Ctrl)) != hipSuccess) rc = TL_E_NOMEM;
    if (rc == TL_OK && hipMemset(c->ctrl, 0, sizeof(Ctrl)) != hipSuccess) rc = TL_E_HIP;
    if (rc == TL_OK && hipHostMalloc((void **)&c->ctrl_h, 2 * sizeof(Ctrl), hipHostMallocDefault) != hipSuccess)
        rc = TL_E_NOMEM;
    if (rc == TL_OK && !events(c)) rc = TL_E_HIP;
    if (rc != TL_OK) {
        g_create_err = c->err.empty() ? "tl_create: HIP setup failed" : c->err;
        tl_destroy(c);
        return rc;
    }
    *out = c;
    return TL_OK;
}

static hipEvent_t *events(tl_ctx *c)
{
    if (!c->ev[0] && hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming) != hipSuccess) return nullptr;
    if (!c->ev[1] && hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming) != hipSuccess) return nullptr;
    return c->ev;
}

extern "C" void tl_destroy(tl_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    unet_destroy(c);
    void *bufs[] = {c->u_alt, c->sd_alt, c->r_alt1, c->r_alt2, c->p_alt, c->p32b, c->guess, c->gcell, c->gc_range, c->p32, c->r32, c->x32,
                    c->r32b, c->part_f, c->partials, c->ctrl, c->hist_a, c->hist_b, c->states_d};
    for (int i = 0; i < TL_NFIELDS; ++i)
        if (c->f[i]) (void)hipFree(c->f[i]);
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->ctrl_h) (void)hipHostFree(c->ctrl_h);
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    for (auto &gb : c->graphs) {
        if (gb.exec) (void)hipGraphExecDestroy(gb.exec);
        if (gb.graph) (void)hipGraphDestroy(gb.graph);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *tl_last_error(const tl_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

extern "C" int tl_dims(const tl_ctx *c, int *X, int *Y)
{
    if (!c) return TL_E_INVALID;
    *X = c->g.X;
    *Y = c->g.Y;
    return TL_OK;
}

// ---------------------------------------------------------------- fields

static double *field_ptr(tl_ctx *c, int id) { return (id >= 0 && id < TL_NFIELDS) ? c->f[id] : nullptr; }

static int copy_in(tl_ctx *c, double *dev, const double *host)
{
    const Geo &g = c->g;
    HIPCHK(c, hipMemcpy2DAsync(dev + (XCOL - g.hd), (size_t)g.pitch * sizeof(double), host,
                               (size_t)g.X * sizeof(double), (size_t)g.X * sizeof(double), (size_t)g.Y,
                               hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

static int copy_out(tl_ctx *c, const double *dev, double *host)
{
    const Geo &g = c->g;
    HIPCHK(c, hipMemcpy2DAsync(host, (size_t)g.X * sizeof(double), dev + (XCOL - g.hd),
                               (size_t)g.pitch * sizeof(double), (size_t)g.X * sizeof(double), (size_t)g.Y,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

extern "C" int tl_set_field(tl_ctx *c, int id, const double *host)
{
    if (!c || !host) return TL_E_INVALID;
    double *d = field_ptr(c, id);
    if (!d) {
        c->err = "tl_set_field: invalid field id";
        return TL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    return copy_in(c, d, host);
}

extern "C" int tl_get_field(tl_ctx *c, int id, double *host)
{
    if (!c || !host) return TL_E_INVALID;
    double *d = field_ptr(c, id);
    if (!d) {
        c->err = "tl_get_field: invalid field id";
        return TL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    return copy_out(c, d, host);
}

extern "C" int tl_set_fields(tl_ctx *c, const double *density, const double *energy)
{
    if (!c) return TL_E_INVALID;
    RETIF(tl_set_field(c, TL_F_DENSITY, density));
    return tl_set_field(c, TL_F_ENERGY, energy);
}

extern "C" int tl_set_guess(tl_ctx *c, const double *host)
{
    if (!c || !host) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->guess) RETIF(alloc_d(c, &c->guess, c->g.elems));
    return copy_in(c, c->guess, host);
}

extern "C" int tl_halo_update(tl_ctx *c, unsigned mask, int depth)
{
    // a reflection of depth d reads d interior cells next to each face
    if (!c || depth < 0 || depth > c->g.hd || depth > c->g.nx || depth > c->g.ny) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_halo(c, mask, depth, nullptr, 0));
    return TL_OK;
}

extern "C" int tl_initialise(tl_ctx *c, const tl_state *states, int nstates)
{
    if (!c || !states || nstates < 1 || nstates > TL_MAX_STATES) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->states_d) HIPCHK(c, hipMalloc((void **)&c->states_d, sizeof(tl_state) * TL_MAX_STATES));
    HIPCHK(c, hipMemcpyAsync(c->states_d, states, sizeof(tl_state) * (size_t)nstates, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, launch_set_state(c, c->states_d, nstates));
    // initialiseapp!, TeaLeaf.jl:38-41
    HIPCHK(c, launch_halo(c, TL_X_DENSITY | TL_X_ENERGY0 | TL_X_ENERGY, 1, nullptr, 0));
    HIPCHK(c, launch_copy_energy(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

// ------------------------------------------------------- hipGraph batches

#define TL_GRAPH_CACHE 4
enum GraphKind { GRAPH_CG64 = 0, GRAPH_CG32 = 1, GRAPH_CGF = 2 };

// what a captured batch of `kind` holds besides the Ctrl-resident scalars:
// batch length, buffers, FP32 face scales and the partial-sum block counts
static std::vector<uintptr_t> graph_key(const tl_ctx *c, int kind, int k)
{
    uint32_t hx = 0, hy = 0;
    std::memcpy(&hx, &c->hx32, sizeof hx);
    std::memcpy(&hy, &c->hy32, sizeof hy);
    std::vector<uintptr_t> v = {(uintptr_t)kind, (uintptr_t)k, (uintptr_t)c->ctrl, (uintptr_t)c->partials,
                                c->nb_r[0], c->nb_r[1], c->nb_w, (uintptr_t)c->gc_fmt};
    if (kind == GRAPH_CG64) {
        for (const void *q : {(const void *)c->f[TL_F_P], (const void *)c->p_alt, (const void *)c->f[TL_F_R],
                              (const void *)c->f[TL_F_KX], (const void *)c->f[TL_F_KY], (const void *)c->f[TL_F_W],
                              (const void *)(c->cg_u ? c->cg_u : c->f[TL_F_U]), (const void *)c->hist_a,
                              (const void *)c->hist_b})
            v.push_back((uintptr_t)q);
    } else {
        for (const void *q : {(const void *)c->p32, (const void *)c->p32b, (const void *)c->r32,
                              (const void *)c->x32, (const void *)c->gcell})
            v.push_back((uintptr_t)q);
        v.push_back(hx);
        v.push_back(hy);
        if (kind == GRAPH_CGF) {
            v.push_back((uintptr_t)c->r32b);
            v.push_back((uintptr_t)c->part_f);
            v.push_back(c->nb_f[0]);
            v.push_back(c->nb_f[1]);
        }
    }
    return v;
}

static void graph_drop(tl_ctx::GraphBatch &b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b = tl_ctx::GraphBatch{};
}

// The executable graph of k CG iterations of `kind`: node 0 advances
// Ctrl::tg, then PW and UR of each iteration.  Every replay starts at an odd
// iteration, so the ping-pong buffers and partial-sum slots of node i are the
// same in every replay.  Called after a stream batch of at least two
// iterations, so the block counts read at capture are the steady ones.
static int graph_get(tl_ctx *c, int kind, int k, hipGraphExec_t *exec)
{
    const std::vector<uintptr_t> key = graph_key(c, kind, k);
    for (auto &gb : c->graphs)
        if (gb.exec && gb.key == key) {
            *exec = gb.exec;
            return TL_OK;
        }
    if (c->graphs.size() < TL_GRAPH_CACHE) c->graphs.emplace_back();
    tl_ctx::GraphBatch &b = c->graphs.size() < TL_GRAPH_CACHE || !c->graphs.back().exec
                                ? c->graphs.back()
                                : c->graphs[c->graph_next++ % TL_GRAPH_CACHE];
    graph_drop(b);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const int tref = 3;  // an odd iteration >= 2 (t only selects buffers and slots here)
    hipError_t e = launch_graph_base(c, k);
    for (int i = 0; i < k && e == hipSuccess; ++i) {
        const int t = tref + i;
        if (kind == GRAPH_CGF) {
            e = launch_cgf(c, t, tref);
            continue;
        }
        e = kind == GRAPH_CG64 ? launch_cg_pw64(c, t, tref) : launch_cg_w32(c, t, tref);
        if (e == hipSuccess) e = kind == GRAPH_CG64 ? launch_cg_urv64(c, t, tref) : launch_cg_ur32(c, t, tref);
    }
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    b.graph = g;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        graph_drop(b);
        c->err = std::string("hipGraph batch: ") + hipGetErrorString(e);
        return TL_E_HIP;
    }
    b.key = key;
    *exec = b.exec;
    return TL_OK;
}

// ------------------------------------------------------- batch pipeline

// Enqueues `body` (one iteration's kernels) in batches, one batch ahead of
// the host's check, until the device control block reports a stop.  Returns
// with the stream synchronised and *out holding the final control block.
// With a graph kind (TL_LAUNCH_GRAPH) only the first batch is enqueued
// through `body`; it ends at an even iteration, and every later batch is one
// replay of a graph of kmax iterations.  *tlast is the last iteration
// enqueued (the counter `body` advances).
template <class Body>
static int run_until_stop(tl_ctx *c, int kmax, long max_batches, Body &&body, Ctrl *out, int gkind = -1,
                          int *tlast = nullptr)
{
    hipEvent_t *ev = events(c);
    Ctrl *snap = c->ctrl_h;
    const bool graph = gkind >= 0 && tlast;
    int k = 4;
    if (graph && ((*tlast + k) & 1)) ++k;
    hipGraphExec_t exec = nullptr;
    for (long b = 0;; ++b) {
        if (graph && b > 0) {
            if (!exec) {
                RETIF(graph_get(c, gkind, kmax, &exec));
                HIPCHK(c, launch_graph_seed(c, *tlast + 1));
            }
            HIPCHK(c, hipGraphLaunch(exec, c->stream));
            *tlast += kmax;
        } else {
            for (int i = 0; i < k; ++i) HIPCHK(c, body());
        }
        HIPCHK(c, hipMemcpyAsync(&snap[b & 1], c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(ev[b & 1], c->stream));
        if (k < kmax) k *= 2;
        if (b > 0) {
            HIPCHK(c, hipEventSynchronize(ev[(b - 1) & 1]));
            if (snap[(b - 1) & 1].stop_iter != INT_MAX || b > max_batches) {
                HIPCHK(c, hipStreamSynchronize(c->stream));
                *out = snap[b & 1];
                if (out->stop_iter == INT_MAX) {
                    c->err = "solve did not stop within its iteration budget";
                    return TL_E_HIP;
                }
                return TL_OK;
            }
        }
    }
}

static int read_ctrl(tl_ctx *c, Ctrl *out)
{
    HIPCHK(c, hipMemcpyAsync(c->ctrl_h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->ctrl_h[0];
    return TL_OK;
}

static int ensure_history(tl_ctx *c, int n)
{
    if (c->hist_len >= n + 2) return TL_OK;
    if (c->hist_a) (void)hipFree(c->hist_a);
    if (c->hist_b) (void)hipFree(c->hist_b);
    c->hist_a = c->hist_b = nullptr;
    c->hist_len = 0;  // nothing held until both allocations succeed
    RETIF(alloc_d(c, &c->hist_a, (size_t)n + 2));
    RETIF(alloc_d(c, &c->hist_b, (size_t)n + 2));
    c->hist_len = n + 2;
    return TL_OK;
}

// The mixed-precision solver's arrays, allocated on its first use.  Each one
// is allocated only while missing, so a call after a failed allocation
// retries the missing ones instead of running with a null array.
static int ensure_fp32(tl_ctx *c)
{
    const size_t n = c->g.elems;
    if (!c->gcell && hipMalloc((void **)&c->gcell, n * sizeof(tl_gc_t)) != hipSuccess) c->gcell = nullptr;
    if (!c->gc_range && hipMalloc((void **)&c->gc_range, 2 * TL_RANGE_BLOCKS * sizeof(float)) != hipSuccess)
        c->gc_range = nullptr;
    if (!c->gcell || !c->gc_range) {
        c->err = "device allocation failed";
        return TL_E_NOMEM;
    }
    if (!c->p32) RETIF(alloc_f(c, &c->p32, n));
    if (!c->p32b) RETIF(alloc_f(c, &c->p32b, n));
    if (!c->r32) RETIF(alloc_f(c, &c->r32, n));
    if (!c->x32) RETIF(alloc_f(c, &c->x32, n));
    return TL_OK;
}

// the fused FP32 kernel's second residual buffer and its per-block sums
static int ensure_fused(tl_ctx *c)
{
    if (!c->r32b) RETIF(alloc_f(c, &c->r32b, c->g.elems));
    if (!c->part_f) RETIF(alloc_d(c, &c->part_f, (size_t)8 * c->g.nbmax));
    return TL_OK;
}

static bool fused_fits(const tl_ctx *c) { return c->g.nx % 4 == 0; }

static int status_of(tl_ctx *c, const Ctrl &k)
{
    if (k.status == TL_E_BREAKDOWN || k.why == WHY_BREAKDOWN) {
        c->err = "solver breakdown: non-positive p.Ap or non-finite residual";
        return TL_E_BREAKDOWN;
    }
    return TL_OK;
}

// ------------------------------------------------------ host scalar work

// Extrema of the eigenvalues of the symmetric tridiagonal (d, e) by Sturm
// sequence bisection (count of eigenvalues below x from the signs of the
// LDL^T pivots), to full double precision.
static int sturm_count(const std::vector<double> &d, const std::vector<double> &e, double x)
{
    int cnt = 0;
    double q = 1.0;
    const size_t n = d.size();
    for (size_t i = 0; i < n; ++i) {
        const double e2 = i ? e[i - 1] * e[i - 1] : 0.0;
        q = d[i] - x - (i ? e2 / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) ++cnt;
    }
    return cnt;
}

static double bisect_kth(const std::vector<double> &d, const std::vector<double> &e, int kth, double lo, double hi)
{
    // smallest x with count(x) > kth  (0-based kth eigenvalue)
    for (int it = 0; it < 200 && hi - lo > 0; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(d, e, mid) > kth)
            hi = mid;
        else
            lo = mid;
    }
    return 0.5 * (lo + hi);
}

extern "C" int tl_eigen_bounds(const double *alpha, const double *beta, int n, double *eigmin, double *eigmax)
{
    // eigenvalues!, kernels.jl:19-51: Lanczos tridiagonal of n CG steps
    if (n < 1) return TL_E_EIGEN;
    std::vector<double> d(n), e(n > 1 ? n - 1 : 0);
    for (int i = 0; i < n; ++i) {
        d[i] = 1.0 / alpha[i];
        if (i > 0) d[i] += beta[i - 1] / alpha[i - 1];
        if (i < n - 1) e[i] = std::sqrt(beta[i]) / alpha[i];
    }
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(d[i])) return TL_E_EIGEN;
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double pad = 1e-12 * std::max(std::fabs(lo), std::fabs(hi)) + 1e-300;
    lo -= pad;
    hi += pad;
    const double mn = bisect_kth(d, e, 0, lo, hi);
    const double mx = bisect_kth(d, e, n - 1, lo, hi);
    if (mn < 0 || mx < 0) return TL_E_EIGEN;  // kernels.jl:41-43
    *eigmin = mn * 0.95;
    *eigmax = mx * 1.05;
    return TL_OK;
}

extern "C" void tl_cheby_coef(double eigmin, double eigmax, int n, double *alpha, double *beta, double *theta)
{
    // Cheby.coef!, Cheby.jl:121-135
    const double th = (eigmax + eigmin) / 2.0;
    const double delta = (eigmax - eigmin) / 2.0;
    const double sigma = th / delta;
    double rho_old = 1.0 / sigma;
    for (int i = 0; i < n; ++i) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        alpha[i] = rho_new * rho_old;
        beta[i] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
    *theta = th;
}

extern "C" int tl_cheby_calciter(double eigmin, double eigmax, double error, double bb)
{
    // Cheby.calciter, Cheby.jl:109-118 (round to nearest, ties to even)
    const double connum = eigmax / eigmin;
    const double ita = 2.220446049250313e-16 * bb / (4.0 * error);
    const double gamma = (std::sqrt(connum) - 1.0) / (std::sqrt(connum) + 1.0);
    const double v = (std::log(ita) / std::log(10.0)) / (2.0 * (std::log(gamma) / std::log(10.0)));
    if (!std::isfinite(v)) return -1;
    return (int)std::nearbyint(v);
}

// --------------------------------------------------------------- solvers

extern "C" int tl_check_solution(const tl_solve_params *p, const tl_solve_stats *st, double bb, double u0sq,
                                 double resid2)
{
    if (!p || !st) return TL_E_INVALID;
    if (!(u0sq >= 0.0) || !(resid2 >= 0.0)) return TL_E_CHECK;  // negative or NaN sums
    // the solve's own sum u0^2 and the recomputed one come from the same field
    if (bb >= 0.0 && std::fabs(bb - u0sq) > 1e-9 * std::max(bb, u0sq)) return TL_E_CHECK;
    const bool cg_family = p->solver == TL_CG || p->solver == TL_CHEBY || p->solver == TL_PPCG;
    const bool relative = p->stop_rule == TL_STOP_RELATIVE && (p->solver == TL_CG || p->precision == TL_FP32_IR);
    const double rel_ok = p->rel_tol * std::sqrt(u0sq) * (1.0 + 1e-6);
    if (st->converged && relative && std::sqrt(resid2) > rel_ok) return TL_E_CHECK;
    if (st->converged && cg_family && st->iterations == 0) {
        // an exact start: the residual the solve found to be zero must be so
        const double r = std::sqrt(resid2);
        if (!(relative ? r <= rel_ok : r < p->eps)) return TL_E_CHECK;
    }
    return TL_OK;
}

// FP64 relative residual of the returned u, recomputed by separate kernels,
// and the check of the solve's convergence claim against it
static int finish_stats(tl_ctx *c, const tl_solve_params *p, tl_solve_stats *st)
{
    Ctrl k0;
    RETIF(read_ctrl(c, &k0));  // the solve's own scalars (bb = sum u0^2 of CG-init)
    const double bb = p->solver == TL_JACOBI && p->precision != TL_FP32_IR ? -1.0 : k0.bb;
    HIPCHK(c, launch_resid_norm_slot(c, c->f[TL_F_U0], c->f[TL_F_U], 2));
    HIPCHK(c, launch_sumsq_slot(c, c->f[TL_F_U0], 0));
    Ctrl k;
    RETIF(read_ctrl(c, &k));
    st->rel_resid = k.sums[0] > 0 ? std::sqrt(k.sums[2] / k.sums[0]) : std::sqrt(k.sums[2]);
    if (tl_check_solution(p, st, bb, k.sums[0], k.sums[2]) != TL_OK) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "convergence check failed: %d iterations, converged %d, recomputed residual %.6e "
                      "(relative %.6e), sum u0^2 %.17g by the solve and %.17g recomputed",
                      st->iterations, st->converged, std::sqrt(k.sums[2]), st->rel_resid, bb, k.sums[0]);
        c->err = msg;
        st->converged = 0;
        return TL_E_CHECK;
    }
    return TL_OK;
}

// Warm start: put the candidate guess in c->guess and keep it only if its
// residual is smaller than that of u = energy*density (the fallback).
// The start of a warm-started solve: the U-Net prediction (TL_WARM_LEARNED)
// or the caller's guess (TL_WARM_GUESS) is used only if its residual
// ||u0 - A x|| is below that of the previous-step start u; otherwise the
// solve starts from u and stats.fallback = 1.
static int choose_guess(tl_ctx *c, const tl_solve_params *p, double rx, double ry, const double **guess,
                        tl_solve_stats *st)
{
    *guess = nullptr;
    st->start = TL_START_PREVIOUS;
    if (p->warm_start == TL_WARM_NONE) return TL_OK;
    if (p->warm_start == TL_WARM_LEARNED && p->warm_policy == TL_WARM_POLICY_BACKOFF) {
        const long s = c->warm_solves++;
        if (s < c->warm_next) return TL_OK;  // waiting after a rejection: start from u
    }
    const auto t0 = std::chrono::steady_clock::now();
    st->warm_attempted = 1;
    if (p->warm_start == TL_WARM_LEARNED) {
        if (!c->guess) RETIF(alloc_d(c, &c->guess, c->g.elems));
        RETIF(unet_predict_guess(c, rx, ry, p->coefficient, c->guess));
    } else if (p->warm_start == TL_WARM_GUESS) {
        if (!c->guess) {
            c->err = "TL_WARM_GUESS without tl_set_guess";
            return TL_E_STATE;
        }
    } else {
        return TL_E_INVALID;
    }
    HIPCHK(c, launch_resid_norm_slot(c, c->f[TL_F_U], c->guess, 0));
    HIPCHK(c, launch_resid_norm_slot(c, c->f[TL_F_U], c->f[TL_F_U], 1));
    Ctrl k;
    RETIF(read_ctrl(c, &k));
    const bool better = std::isfinite(k.sums[0]) && k.sums[0] < k.sums[1];
    if (better) {
        *guess = c->guess;
        st->start = TL_START_WARM;
    } else {
        st->fallback = 1;
    }
    if (p->warm_start == TL_WARM_LEARNED && p->warm_policy == TL_WARM_POLICY_BACKOFF) {
        // accepted: attempt again next solve; rejected: wait 2, 4, 8, then 16 solves
        c->warm_gap = better ? 1 : std::min(2 * c->warm_gap, 16);
        c->warm_next = c->warm_solves - 1 + c->warm_gap;
    }
    st->warm_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return TL_OK;
}

static void fill_stats(const Ctrl &k, tl_solve_stats *st)
{
    st->error = k.err;
    st->converged = k.why == WHY_CONVERGED;
}

static hipError_t cg64_iteration(tl_ctx *c, int t)
{
    hipError_t e = launch_cg_w64(c, t);
    if (e == hipSuccess) e = launch_cg_ur64(c, t);
    if (e == hipSuccess) e = launch_cg_p64(c, t);
    return e;
}

// CG.solve!, CG.jl:7-29 (FP64)
static int solve_cg64(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    RETIF(ensure_history(c, p->max_iters));
    HIPCHK(c, launch_init_coef(c, rx, ry, p->coefficient, INIT_CG));
    const int mode = p->stop_rule == TL_STOP_RELATIVE ? MODE_CG_REL : MODE_CG_REF;
    HIPCHK(c, launch_init_ctrl(c, mode, p->max_iters, p->eps, p->rel_tol, 0.0, p->presteps, p->error_switch,
                               p->epslim));
    const double *guess = nullptr;
    RETIF(choose_guess(c, p, rx, ry, &guess, st));
    HIPCHK(c, launch_cg_init_resid(c, guess));
    // Under the relative rule the iterations add alpha p to a correction d
    // (u_alt, free during CG) instead of to u, and d is added to u at every
    // true-residual check ("reliable updates" with group-wise solution
    // updates): the rounding of the updates is then relative to the small
    // correction, not to u, so the true residual can reach tolerances that
    // CG updating u directly stalls above (measured at 8192^2 from the 69th
    // timestep of a random-rectangle problem: 2-3e-9 against 1e-10).
    const bool group = mode == MODE_CG_REL;
    if (group) HIPCHK(c, hipMemsetAsync(c->u_alt, 0, c->g.elems * sizeof(double), c->stream));
    c->cg_u = group ? c->u_alt : nullptr;
    Ctrl k;
    int t = 0;
    for (;;) {
        // fused iterations: PW(t) = [end of t-1: beta, stop test, p = r + beta p] + w = A p; UR(t)
        RETIF(run_until_stop(c, 64, (long)p->max_iters + 16,
                             [&]() -> hipError_t {
                                 ++t;
                                 hipError_t e = launch_cg_pw64(c, t);
                                 return e == hipSuccess ? launch_cg_urv64(c, t) : e;
                             },
                             &k, p->launch == TL_LAUNCH_GRAPH ? GRAPH_CG64 : -1, &t));
        if (mode != MODE_CG_REL || k.why != WHY_CONVERGED) break;
        // the relative rule is met by the recursive residual: confirm it on the
        // true one, and if rounding has let the two drift apart, continue CG
        // from the true residual (keeping the search direction)
        t = k.stop_iter;
        HIPCHK(c, launch_acc64(c, c->f[TL_F_U], c->u_alt));
        HIPCHK(c, launch_cg64_true_residual(c, t));
        RETIF(read_ctrl(c, &k));
        if (k.stop_iter != INT_MAX) break;
        ++st->refreshes;
    }
    if (group) HIPCHK(c, launch_acc64(c, c->f[TL_F_U], c->u_alt));  // pending updates (max_iters, breakdown)
    c->cg_u = nullptr;  // other callers of the UR launcher update u
    const int T = k.stop_iter;
    if (T >= 1) {
        // the last iteration's CG.p!, so p ends as the reference leaves it
        HIPCHK(c, launch_cg_pfinal64(c, T));
        double *pT = ((T - 1) & 1) ? c->p_alt : c->f[TL_F_P];
        if (pT != c->f[TL_F_P]) std::swap(c->f[TL_F_P], c->p_alt);
        RETIF(read_ctrl(c, &k));
    }
    st->init_rel_resid = k.bb > 0 ? std::sqrt(k.sums[0] / k.bb) : 0.0;
    st->iterations = st->cg_iters = T;
    c->last_cg_iters = T;
    fill_stats(k, st);
    HIPCHK(c, launch_halo(c, TL_X_U | TL_X_P, 1, nullptr, 0));
    return status_of(c, k);
}

// the CG phase shared by Cheby and PPCG: CG-init, then CG steps until the
// switch test of Cheby.jl:25-29 / PPCG.jl:25-30 (or convergence / max_iters)
static int cg_phase(tl_ctx *c, const tl_solve_params *p, double rx, double ry, Ctrl *k, tl_solve_stats *st)
{
    RETIF(ensure_history(c, p->max_iters));
    HIPCHK(c, launch_init_coef(c, rx, ry, p->coefficient, INIT_CG));
    HIPCHK(c, launch_init_ctrl(c, MODE_SWITCH, p->max_iters, p->eps, p->rel_tol, 0.0, p->presteps,
                               p->error_switch, p->epslim));
    HIPCHK(c, launch_cg_init_resid(c, nullptr));
    int t = 0;
    RETIF(run_until_stop(c, 32, (long)p->max_iters + 16, [&]() { return cg64_iteration(c, ++t); }, k));
    st->init_rel_resid = k->bb > 0 ? std::sqrt(k->sums[0] / k->bb) : 0.0;
    st->cg_iters = k->stop_iter;
    st->iterations = k->stop_iter;
    c->last_cg_iters = k->stop_iter;
    fill_stats(*k, st);
    return status_of(c, *k);
}

static int eigen_from_history(tl_ctx *c, int cgiters, double *eigmin, double *eigmax)
{
    std::vector<double> a(cgiters), b(cgiters);
    HIPCHK(c, hipMemcpyAsync(a.data(), c->hist_a, sizeof(double) * cgiters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.data(), c->hist_b, sizeof(double) * cgiters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int rc = tl_eigen_bounds(a.data(), b.data(), cgiters, eigmin, eigmax);
    if (rc != TL_OK) c->err = "eigenvalue estimate failed (negative or non-finite)";
    return rc;
}

// Cheby.solve!, Cheby.jl:10-61 (intended form)
static int solve_cheby(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    Ctrl k;
    RETIF(cg_phase(c, p, rx, ry, &k, st));
    if (k.why != WHY_SWITCH) {
        HIPCHK(c, launch_halo(c, TL_X_U | TL_X_P, 1, nullptr, 0));
        return TL_OK;
    }
    const int cgiters = k.stop_iter;
    const int maxiters = p->max_iters;
    double eigmin, eigmax, theta;
    RETIF(eigen_from_history(c, cgiters, &eigmin, &eigmax));
    st->eigmin = eigmin;
    st->eigmax = eigmax;
    const int ncoef = maxiters - cgiters;
    std::vector<double> ca(maxiters + 2, 0.0), cb(maxiters + 2, 0.0);  // 1-based, zero past ncoef
    {
        std::vector<double> ta(std::max(ncoef, 1)), tb(std::max(ncoef, 1));
        tl_cheby_coef(eigmin, eigmax, ncoef, ta.data(), tb.data(), &theta);
        for (int i = 0; i < ncoef; ++i) {
            ca[i + 1] = ta[i];
            cb[i + 1] = tb[i];
        }
    }
    HIPCHK(c, launch_set_mode(c, MODE_JACOBI, maxiters));
    double *ucur = c->f[TL_F_U], *unext = c->u_alt;
    double *pcur = c->f[TL_F_P], *pnext = c->p_alt;  // p moves only in two-iteration launches
    const bool pairs = p->cheby_pairs == TL_CHEBY_PAIRS_AUTO;
    // u (and p) move between two buffers and either may end up as the
    // field: the cells no kernel writes (the outer halo) start equal in both
    HIPCHK(c, launch_copy_full(c, ucur, unext));
    if (pairs) HIPCHK(c, launch_copy_full(c, pcur, pnext));
    // Cheby.init! and the first main step (Cheby.jl:40-45); iteration t = tt
    int tt = cgiters + 1;
    int chebyiters = 1;
    HIPCHK(c, launch_cheby_init(c, theta, ucur, unext));
    std::swap(ucur, unext);
    HIPCHK(c, launch_cheby_iter(c, ca[2], cb[2], ucur, unext, pcur, 1, tt));
    std::swap(ucur, unext);
    RETIF(read_ctrl(c, &k));
    const int estiter = tl_cheby_calciter(eigmin, eigmax, k.err, k.bb);
    st->est_iters = estiter;
    if (k.stop_iter == INT_MAX) {
        // main loop (Cheby.jl:46-53); every launch swaps the u buffers.  Two
        // iterations per launch (TL_CHEBY_PAIRS_AUTO) unless the first needs
        // the residual norm, so that a stop can only follow a launch's last
        // iteration
        while (tt < maxiters) {
            int nb = 0;
            int norm = 0;
            while (tt < maxiters && nb < 64) {
                const bool norm1 = (chebyiters + 1 >= estiter) && ((tt + 2) % 10 == 0);
                if (pairs && !norm1 && tt + 2 <= maxiters) {
                    norm = (chebyiters + 2 >= estiter) && ((tt + 3) % 10 == 0);
                    HIPCHK(c, launch_cheby_iter2(c, ca[chebyiters + 2], cb[chebyiters + 2], ca[chebyiters + 3],
                                                 cb[chebyiters + 3], ucur, unext, pcur, pnext, norm, tt + 1));
                    std::swap(pcur, pnext);
                    tt += 2;
                    chebyiters += 2;
                    nb += 2;
                } else {
                    ++tt;
                    ++chebyiters;
                    norm = (chebyiters >= estiter) && ((tt + 1) % 10 == 0);
                    HIPCHK(c, launch_cheby_iter(c, ca[chebyiters + 1], cb[chebyiters + 1], ucur, unext, pcur, norm,
                                                tt));
                    ++nb;
                }
                std::swap(ucur, unext);
                if (norm) break;
            }
            if (norm) {
                RETIF(read_ctrl(c, &k));
                if (k.stop_iter != INT_MAX) break;
            }
        }
        RETIF(read_ctrl(c, &k));
        if (pairs && k.stop_iter != INT_MAX && k.stop_iter != tt) {
            // stops follow norm iterations, which end a launch
            c->err = "Chebyshev: stop inside a two-iteration launch";
            return TL_E_STATE;
        }
        if (k.stop_iter != INT_MAX && ((tt - k.stop_iter) & 1)) std::swap(ucur, unext);
        if (k.stop_iter != INT_MAX) {
            chebyiters -= tt - k.stop_iter;
            tt = k.stop_iter;
        }
    }
    if (ucur != c->f[TL_F_U]) std::swap(c->f[TL_F_U], c->u_alt);
    if (pcur != c->f[TL_F_P]) std::swap(c->f[TL_F_P], c->p_alt);
    st->iterations = tt;
    st->cheby_iters = chebyiters;
    st->error = k.err;
    st->converged = std::fabs(k.err) < p->eps;
    HIPCHK(c, launch_halo(c, TL_X_U | TL_X_P, 1, nullptr, 0));
    return TL_OK;
}

// PPCG.solve!, PPCG.jl:9-55 (intended form)
static int solve_ppcg(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    Ctrl k;
    RETIF(cg_phase(c, p, rx, ry, &k, st));
    if (k.why != WHY_SWITCH) {
        HIPCHK(c, launch_halo(c, TL_X_U | TL_X_P, 1, nullptr, 0));
        return TL_OK;
    }
    const int cgiters = k.stop_iter;
    const int inner = p->ppcg_inner_steps;
    double eigmin, eigmax, theta;
    RETIF(eigen_from_history(c, cgiters, &eigmin, &eigmax));
    st->eigmin = eigmin;
    st->eigmax = eigmax;
    std::vector<double> ca(std::max(inner, 1)), cb(std::max(inner, 1));
    tl_cheby_coef(eigmin, eigmax, inner, ca.data(), cb.data(), &theta);
    // PPCG.init!(chunk, settings), PPCG.jl:58-61: r = u0 - A u
    HIPCHK(c, launch_residual(c));
    HIPCHK(c, launch_set_mode(c, MODE_PPCG, p->max_iters));
    double *sd0 = c->f[TL_F_SD], *sd1 = c->sd_alt;
    // two inner steps per sweep (r through two spare buffers, back in r's
    // own at the end of every outer iteration), an odd last step alone
    const bool pairs = p->ppcg_pairs == TL_PPCG_PAIRS_AUTO && inner >= 4;
    if (pairs) {
        if (!c->r_alt1) RETIF(alloc_d(c, &c->r_alt1, c->g.elems));
        if (!c->r_alt2) RETIF(alloc_d(c, &c->r_alt2, c->g.elems));
    }
    const int sd_moves = pairs ? inner / 2 + (inner & 1) : inner;  // sweeps per outer iteration
    int t = cgiters;
    RETIF(run_until_stop(c, 16, (long)p->max_iters + 16,
                         [&]() -> hipError_t {
                             ++t;
                             hipError_t e = launch_cg_w64(c, t);
                             if (e == hipSuccess) e = launch_ppcg_ur_sd(c, theta, t);
                             double *a = sd0, *b = sd1;
                             int pp = 0;
                             if (pairs) {
                                 const int np = inner / 2;
                                 const double *rs = c->f[TL_F_R];
                                 for (int q = 0; q < np && e == hipSuccess; ++q, pp += 2) {
                                     double *rd = q == np - 1 ? c->f[TL_F_R] : (q & 1) ? c->r_alt2 : c->r_alt1;
                                     e = launch_ppcg_inner2(c, ca[pp], cb[pp], ca[pp + 1], cb[pp + 1], a, b, rs, rd,
                                                            pp + 2 == inner, t);
                                     std::swap(a, b);
                                     rs = rd;
                                 }
                             }
                             for (; pp < inner && e == hipSuccess; ++pp) {
                                 e = launch_ppcg_inner(c, ca[pp], cb[pp], a, b, pp == inner - 1, t);
                                 std::swap(a, b);
                             }
                             if (inner == 0 && e == hipSuccess) e = launch_sumsq_rr(c, c->f[TL_F_R], t);
                             if (e == hipSuccess) e = launch_cg_p64(c, t);
                             return e;
                         },
                         &k));
    if (sd_moves & 1) std::swap(c->f[TL_F_SD], c->sd_alt);
    st->cheby_iters = k.stop_iter - cgiters;
    st->iterations = k.stop_iter;
    fill_stats(k, st);
    HIPCHK(c, launch_halo(c, TL_X_U | TL_X_P | TL_X_SD, 1, nullptr, 0));
    return status_of(c, k);
}

// Jacobi.driver!, Jacobi.jl:7-31 (intended form)
static int solve_jacobi(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    HIPCHK(c, launch_init_coef(c, rx, ry, p->coefficient, INIT_JACOBI));
    HIPCHK(c, launch_init_ctrl(c, MODE_JACOBI, p->max_iters, p->eps, p->rel_tol, 0.0, 0, 0, p->epslim));
    double *ucur = c->f[TL_F_U], *unext = c->u_alt;
    int tt = 0;
    Ctrl k;
    RETIF(run_until_stop(c, 50, (long)p->max_iters + 16,
                         [&]() -> hipError_t {
                             if (tt >= p->max_iters) return hipSuccess;
                             ++tt;
                             const int every = (tt % 50) == 0;
                             hipError_t e = launch_jacobi(c, ucur, unext, tt, every);
                             if (every && e == hipSuccess) e = launch_jacobi_residual(c, unext, tt);
                             std::swap(ucur, unext);
                             return e;
                         },
                         &k));
    // tt launches were made; the last one that ran was k.stop_iter
    if ((tt - k.stop_iter) & 1) std::swap(ucur, unext);
    if (ucur != c->f[TL_F_U]) std::swap(c->f[TL_F_U], c->u_alt);
    st->iterations = k.stop_iter;
    fill_stats(k, st);
    HIPCHK(c, launch_halo(c, TL_X_U, 1, nullptr, 0));
    return status_of(c, k);
}

// one fused FP32 iteration: PW(t) (end of t-1 + w = A p) and UR(t)
static hipError_t cg32_iteration(tl_ctx *c, int t)
{
    hipError_t e = launch_cg_w32(c, t);
    if (e == hipSuccess) e = launch_cg_ur32(c, t);
    return e;
}

// FP32 inner CG with FP64 residual replacement ("reliable updates"): the
// FP32 recurrence runs on the correction x; when its residual has fallen by
// ir_delta (or reaches the tolerance) the FP64 solution is accumulated,
// u += x, and the FP64 residual r = u0 - A u replaces the FP32 one, keeping
// the search direction.  Converged when the FP64 residual meets rel_tol.
static int solve_cg_ir(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    RETIF(ensure_fp32(c));
    HIPCHK(c, launch_init_coef(c, rx, ry, p->coefficient, INIT_CG));
    HIPCHK(c, launch_ir_prepare(c, rx, ry, p->coefficient));
    const double delta = p->ir_delta > 0 && p->ir_delta < 1 ? p->ir_delta : 0.1;
    HIPCHK(c, launch_init_ctrl(c, MODE_IR, p->max_iters, p->eps, p->rel_tol, delta, 0, 0, p->epslim));
    // pairs of iterations (k_cgp) run as stream launches; with TL_LAUNCH_GRAPH
    // the automatic choice takes the fused kernel
    const bool pairs = p->fp32_kernels == TL_FP32_KERNELS_PAIRS ||
                       (p->fp32_kernels == TL_FP32_KERNELS_AUTO && TL_CGP_AUTO && fused_fits(c) &&
                        p->launch != TL_LAUNCH_GRAPH);
    const bool fused = !pairs && (p->fp32_kernels == TL_FP32_KERNELS_FUSED ||
                                  (p->fp32_kernels == TL_FP32_KERNELS_AUTO && fused_fits(c)));
    if (fused || pairs) RETIF(ensure_fused(c));
    const double *guess = nullptr;
    RETIF(choose_guess(c, p, rx, ry, &guess, st));
    // u0 = b = energy*density; u = guess or b; bb, tol_abs
    HIPCHK(c, launch_cg_init_resid(c, guess));
    HIPCHK(c, hipMemsetAsync(c->x32, 0, c->g.elems * sizeof(float), c->stream));
    int t = 0;  // last iteration enqueued
    // r = u0 - A u (FP64, stored in FP32) and p = r; fused: the four sums of
    // iteration 1 (r and p in the buffers of iteration 0)
    HIPCHK(c, launch_ir_residual(c, 1, t + 1));
    if (fused) HIPCHK(c, launch_cgf_dots(c, 0));
    int tfirst = 1;  // pairs: the first pair after a start or a refresh
    Ctrl k;
    RETIF(read_ctrl(c, &k));
    st->init_rel_resid = k.bb > 0 ? std::sqrt(k.sums[0] / k.bb) : 0.0;
    int refreshes = 0;
    while (k.stop_iter == INT_MAX || k.why == WHY_REFRESH) {
        if (k.stop_iter != INT_MAX && k.why == WHY_REFRESH) {
            t = k.stop_iter;
            if (pairs) {
                // u += x (complete after every pair), the FP64 residual into
                // r(t)'s buffer, p(t) kept
                HIPCHK(c, launch_ir_accumulate(c, t, nullptr));
                HIPCHK(c, launch_ir_residual(c, 0, t + 1, cgp_r(c, t)));
                tfirst = t + 1;
            } else if (fused) {
                // u += x (with the pending update of an odd t), the FP64
                // residual into r(t)'s buffer, p(t) kept, and the four sums
                HIPCHK(c, launch_ir_accumulate(c, t, cgf_p(c, t)));
                HIPCHK(c, launch_ir_residual(c, 0, t + 1, cgf_r(c, t)));
                HIPCHK(c, launch_cgf_dots(c, t));
            } else {
                HIPCHK(c, launch_ir_accumulate(c, t));
                HIPCHK(c, launch_ir_residual(c, 0, t + 1));
            }
            ++refreshes;
            RETIF(read_ctrl(c, &k));
            continue;
        }
        if (pairs)
            RETIF(run_until_stop(c, 32, (long)p->max_iters / 2 + 16,
                                 [&]() {
                                     const hipError_t e = launch_cgp(c, t + 1, t + 1 == tfirst);
                                     t += 2;
                                     return e;
                                 },
                                 &k));
        else if (fused)
            RETIF(run_until_stop(c, 64, (long)p->max_iters + 16, [&]() { return launch_cgf(c, ++t); }, &k,
                                 p->launch == TL_LAUNCH_GRAPH ? GRAPH_CGF : -1, &t));
        else
            RETIF(run_until_stop(c, 64, (long)p->max_iters + 16, [&]() { return cg32_iteration(c, ++t); }, &k,
                                 p->launch == TL_LAUNCH_GRAPH ? GRAPH_CG32 : -1, &t));
    }
    if (k.why == WHY_MAXITERS)  // keep the partial correction
        HIPCHK(c, launch_ir_accumulate(c, k.stop_iter, fused ? cgf_p(c, k.stop_iter) : nullptr));
    st->inner_iters = k.iters;
    st->iterations = k.iters;
    st->cg_iters = k.iters;
    st->refreshes = refreshes;
    st->coef_format = !TL_GC16 ? TL_COEF_FP32 : c->gc_fmt == GC_F16 ? TL_COEF_FP16 : TL_COEF_BF16;
    fill_stats(k, st);
    HIPCHK(c, launch_halo(c, TL_X_U, 1, nullptr, 0));
    return status_of(c, k);
}

// collect the durations of the profiled launches recorded so far
static int prof_harvest(tl_ctx *c)
{
    if (c->prof_used == 0) return TL_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i + 1 < c->prof_used; i += 2) {
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->prof_ev[i], c->prof_ev[i + 1]));
        c->prof_ms[c->prof_tag[i / 2]].push_back(ms);
    }
    c->prof_used = 0;
    return TL_OK;
}

extern "C" int tl_profile_begin(tl_ctx *c, unsigned kernel_mask, int period)
{
    if (!c || (kernel_mask >> TL_K_COUNT) || period < 1) return TL_E_INVALID;
    RETIF(prof_harvest(c));
    c->prof_mask = kernel_mask;
    c->prof_period = (unsigned)period;
    for (auto &n : c->prof_seq) n = 0;
    for (auto &v : c->prof_ms) v.clear();
    return TL_OK;
}

extern "C" int tl_profile_read(tl_ctx *c, int kernel_id, double *durations_ms, int max, int *count)
{
    if (!c || !count || kernel_id < 0 || kernel_id >= TL_K_COUNT) return TL_E_INVALID;
    RETIF(prof_harvest(c));
    const std::vector<double> &v = c->prof_ms[kernel_id];
    *count = (int)v.size();
    if (durations_ms)
        for (int i = 0; i < *count && i < max; ++i) durations_ms[i] = v[i];
    return TL_OK;
}

extern "C" int tl_profile_end(tl_ctx *c)
{
    if (!c) return TL_E_INVALID;
    RETIF(prof_harvest(c));
    c->prof_mask = 0;
    for (auto &v : c->prof_ms) v.clear();
    return TL_OK;
}

extern "C" int tl_kernel_repeat(tl_ctx *c, int kernel_id, int reps)
{
    if (!c || reps < 1 || kernel_id < -2 || kernel_id >= TL_K_COUNT) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    if (kernel_id < 0) {
        for (int i = 0; i < reps; ++i)
            HIPCHK(c, kernel_id == -1 ? launch_copy_full(c, c->f[TL_F_U], c->u_alt)
                                      : launch_copy_vec(c, c->f[TL_F_U], c->u_alt));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TL_OK;
    }
    // a live CG state whose stop test never fires: p = r != 0, eps = 0
    RETIF(ensure_history(c, 1 << 20));
    HIPCHK(c, launch_init_coef(c, 1.0, 1.0, 1, INIT_CG));
    HIPCHK(c, launch_init_ctrl(c, MODE_CG_REF, 1 << 30, 0.0, 0.0, 0.1, 0, 0, 0.0));
    HIPCHK(c, launch_cg_init_resid(c, nullptr));
    if (kernel_id == TL_K_CGF32) {
        // `reps` iterations of the fused FP32 CG (after three warm-up ones)
        if (!fused_fits(c)) return TL_E_INVALID;
        RETIF(ensure_fp32(c));
        RETIF(ensure_fused(c));
        HIPCHK(c, launch_ir_prepare(c, 1.0, 1.0, 1));
        HIPCHK(c, launch_ir_residual(c, 1, 1));
        HIPCHK(c, launch_cgf_dots(c, 0));
        for (int t = 1; t <= 3 + reps; ++t) HIPCHK(c, launch_cgf(c, t));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TL_OK;
    }
    if (kernel_id == TL_K_CGP_G32 || kernel_id == TL_K_CGP_U32) {
        // `reps` pairs of FP32 CG iterations (after two warm-up pairs); both
        // sweeps of a pair run whichever of the two is named
        if (!fused_fits(c)) return TL_E_INVALID;
        RETIF(ensure_fp32(c));
        RETIF(ensure_fused(c));
        HIPCHK(c, launch_ir_prepare(c, 1.0, 1.0, 1));
        HIPCHK(c, launch_ir_residual(c, 1, 1));
        for (int i = 0; i < 2 + reps; ++i) HIPCHK(c, launch_cgp(c, 1 + 2 * i, i == 0));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TL_OK;
    }
    if (kernel_id >= TL_K_CHEBY64) {
        // one FP64 solver kernel, `reps` launches on the CG-initialised fields
        // (u, u0, p, r, kx, ky) with fixed coefficients; never stopped
        HIPCHK(c, launch_set_mode(c, MODE_JACOBI, 1 << 30));
        if (kernel_id == TL_K_PPCG2_64 && !c->r_alt1) RETIF(alloc_d(c, &c->r_alt1, c->g.elems));
        double *a = c->f[TL_F_U], *b = c->u_alt, *s0 = c->f[TL_F_SD], *s1 = c->sd_alt;
        double *r0 = c->f[TL_F_R], *r1 = c->r_alt1, *q0 = c->f[TL_F_P], *q1 = c->p_alt;
        for (int i = 0; i < reps; ++i) {
            const int t = 1 + i;
            switch (kernel_id) {
            case TL_K_CHEBY64:
                HIPCHK(c, launch_cheby_iter(c, 0.5, 1e-3, a, b, c->f[TL_F_P], 0, t));
                std::swap(a, b);
                break;
            case TL_K_CHEBY2_64:
                HIPCHK(c, launch_cheby_iter2(c, 0.5, 1e-3, 0.5, 1e-3, a, b, q0, q1, 0, t));
                std::swap(a, b);
                std::swap(q0, q1);
                break;
            case TL_K_PPCG64: HIPCHK(c, launch_ppcg_inner(c, 0.5, 1e-3, s0, s1, 0, t)); std::swap(s0, s1); break;
            case TL_K_PPCG2_64:
                HIPCHK(c, launch_ppcg_inner2(c, 0.5, 1e-3, 0.5, 1e-3, s0, s1, r0, r1, 0, t));
                std::swap(s0, s1);
                std::swap(r0, r1);
                break;
            case TL_K_JACOBI64: HIPCHK(c, launch_jacobi(c, a, b, t, 0)); std::swap(a, b); break;
            default: HIPCHK(c, launch_residual(c)); break;
            }
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TL_OK;
    }
    if (kernel_id >= TL_K_CG_W32) {
        RETIF(ensure_fp32(c));
        HIPCHK(c, launch_ir_prepare(c, 1.0, 1.0, 1));
        HIPCHK(c, launch_ir_residual(c, 1, 1));
    }
    // three live iterations so every partial-sum array and ping-pong buffer
    // is filled, then `reps` further iterations of the same CG (both kernels
    // of each; consecutive iterations alternate the FP32 UR kernel's even and
    // odd forms), or `reps` launches of the FP64 P kernel of iteration 4
    for (int t = 1; t <= 3; ++t) {
        if (kernel_id >= TL_K_CG_W32) {
            HIPCHK(c, cg32_iteration(c, t));
        } else {
            HIPCHK(c, launch_cg_pw64(c, t));
            HIPCHK(c, launch_cg_urv64(c, t));
        }
    }
    for (int i = 0; i < reps; ++i) {
        const int t = 4 + i;
        if (kernel_id >= TL_K_CG_W32) {
            HIPCHK(c, cg32_iteration(c, t));
        } else if (kernel_id == TL_K_CG_P64) {
            HIPCHK(c, launch_cg_p64(c, 4));
        } else {
            HIPCHK(c, launch_cg_pw64(c, t));
            HIPCHK(c, launch_cg_urv64(c, t));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

extern "C" int tl_solve(tl_ctx *c, const tl_solve_params *p, double rx, double ry, tl_solve_stats *st)
{
    if (!c || !p || !st) return TL_E_INVALID;
    std::memset(st, 0, sizeof(*st));
    if (p->coefficient != 1 && p->coefficient != 2) {
        c->err = "Coefficient " + std::to_string(p->coefficient) + " is not valid";  // CG.jl:48-50
        return TL_E_INVALID;
    }
    if (p->max_iters < 1 || p->solver < TL_CG || p->solver > TL_JACOBI || p->ppcg_inner_steps < 0) {
        c->err = "tl_solve: invalid parameters";
        return TL_E_INVALID;
    }
    if (p->precision == TL_FP32_IR && (p->solver != TL_CG || p->stop_rule != TL_STOP_RELATIVE)) {
        c->err = "TL_FP32_IR needs the CG solver and the relative stop rule";
        return TL_E_INVALID;
    }
    if (p->fp32_kernels < TL_FP32_KERNELS_AUTO || p->fp32_kernels > TL_FP32_KERNELS_PAIRS ||
        (p->precision == TL_FP32_IR &&
         (p->fp32_kernels == TL_FP32_KERNELS_FUSED || p->fp32_kernels == TL_FP32_KERNELS_PAIRS) && !fused_fits(c))) {
        c->err = "tl_solve: fp32_kernels must be TL_FP32_KERNELS_AUTO, _TWO, _FUSED or _PAIRS (fused and pairs need "
                 "nx % 4 == 0)";
        return TL_E_INVALID;
    }
    if (p->launch != TL_LAUNCH_STREAM && p->launch != TL_LAUNCH_GRAPH) {
        c->err = "tl_solve: launch must be TL_LAUNCH_STREAM or TL_LAUNCH_GRAPH";
        return TL_E_INVALID;
    }
    if (p->warm_policy != TL_WARM_POLICY_BACKOFF && p->warm_policy != TL_WARM_POLICY_EVERY) {
        c->err = "tl_solve: warm_policy must be TL_WARM_POLICY_BACKOFF or TL_WARM_POLICY_EVERY";
        return TL_E_INVALID;
    }
    if (p->ppcg_pairs != TL_PPCG_PAIRS_AUTO && p->ppcg_pairs != TL_PPCG_PAIRS_OFF) {
        c->err = "tl_solve: ppcg_pairs must be TL_PPCG_PAIRS_AUTO or TL_PPCG_PAIRS_OFF";
        return TL_E_INVALID;
    }
    if (p->cheby_pairs != TL_CHEBY_PAIRS_AUTO && p->cheby_pairs != TL_CHEBY_PAIRS_OFF) {
        c->err = "tl_solve: cheby_pairs must be TL_CHEBY_PAIRS_AUTO or TL_CHEBY_PAIRS_OFF";
        return TL_E_INVALID;
    }
    if (p->warm_start != TL_WARM_NONE && p->solver != TL_CG) {
        c->err = "warm starts apply to the CG solver";
        return TL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if (p->precision == TL_FP32_IR) {
        rc = solve_cg_ir(c, p, rx, ry, st);
    } else {
        switch (p->solver) {
        case TL_CG: rc = solve_cg64(c, p, rx, ry, st); break;
        case TL_CHEBY: rc = solve_cheby(c, p, rx, ry, st); break;
        case TL_PPCG: rc = solve_ppcg(c, p, rx, ry, st); break;
        default: rc = solve_jacobi(c, p, rx, ry, st); break;
        }
    }
    if (rc != TL_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const auto t1 = std::chrono::steady_clock::now();
    st->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    RETIF(finish_stats(c, p, st));
    RETIF(prof_harvest(c));
    // training data for the learned warm start: (inputs of this step, solution)
    if (c->recording) RETIF(unet_record_step(c, rx, ry, p->coefficient));
    return TL_OK;
}

extern "C" int tl_solve_finished(tl_ctx *c, int check_result)
{
    if (!c) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    // solvefinished!, kernels.jl:166-170
    if (check_result) HIPCHK(c, launch_residual(c));
    HIPCHK(c, launch_finalise(c));
    HIPCHK(c, launch_halo(c, TL_X_ENERGY, 1, nullptr, 0));
    return TL_OK;
}

extern "C" int tl_field_summary(tl_ctx *c, double out[4])
{
    if (!c || !out) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_field_summary(c));
    Ctrl k;
    RETIF(read_ctrl(c, &k));
    for (int i = 0; i < 4; ++i) out[i] = k.sums[i];
    return TL_OK;
}

extern "C" int tl_step(tl_ctx *c, const tl_solve_params *p, double dt, tl_solve_stats *st)
{
    if (!c || !p) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    // diffuse!, TeaLeaf.jl:69-77
    const double rx = dt / (c->dx * c->dx), ry = dt / (c->dy * c->dy);
    HIPCHK(c, launch_halo(c, TL_X_ENERGY | TL_X_DENSITY, 1, nullptr, 0));
    RETIF(tl_solve(c, p, rx, ry, st));
    return tl_solve_finished(c, p->check_result);
}

extern "C" int tl_apply_operator(tl_ctx *c, int field_id, double rx, double ry, int coefficient)
{
    if (!c) return TL_E_INVALID;
    double *src = field_ptr(c, field_id);
    if (!src || field_id == TL_F_W || field_id == TL_F_KX || field_id == TL_F_KY) return TL_E_INVALID;
    if (coefficient != 1 && coefficient != 2) return TL_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    // coefficients from the current density (CG.jl:56-68), fields untouched
    HIPCHK(c, launch_init_coef(c, rx, ry, coefficient, INIT_COEF_ONLY));
    HIPCHK(c, launch_apply(c, src, c->f[TL_F_W]));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

extern "C" int tl_get_cg_history(tl_ctx *c, double *alpha, double *beta, int n)
{
    if (!c || n < 0 || n > c->hist_len) return TL_E_INVALID;
    if (n == 0) return TL_OK;
    HIPCHK(c, hipMemcpyAsync(alpha, c->hist_a, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(beta, c->hist_b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

// ------------------------------------------------------------ deck / run

extern "C" int tl_deck_create(tl_ctx **out, const tl_deck *d, int device)
{
    if (!out || !d) return TL_E_INVALID;
    RETIF(tl_create(out, d->x_cells, d->y_cells, d->halo_depth, d->xmin, d->ymin, d->dx, d->dy, device));
    const int rc = tl_initialise(*out, d->states, d->nstates);
    if (rc != TL_OK) {
        g_create_err = (*out)->err;
        tl_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

extern "C" int tl_run(tl_ctx *c, const tl_deck *d, const tl_solve_params *p, int nsteps, tl_step_cb cb, void *user)
{
    if (!c || !d) return TL_E_INVALID;
    const tl_solve_params *pp = p ? p : &d->params;
    const long long n = nsteps > 0 ? nsteps : d->end_step;
    for (long long t = 1; t <= n; ++t) {
        tl_solve_stats st;
        RETIF(tl_step(c, pp, d->dt_init, &st));
        double s[4];
        RETIF(tl_field_summary(c, s));
        if (cb) cb((int)t, s, &st, user);
    }
    return TL_OK;
}
