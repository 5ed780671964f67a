// tl_kernels.hip -- gfx950 kernels of the TeaLeaf linear-solve hot path.
//
// Launch shape.  A block of 256 threads owns a 256-cell row segment and
// sweeps ty consecutive rows of it; each thread owns one column.  Rows are
// processed in chunks of R = 8 whose loads are all issued before any
// arithmetic (R rows of loads in flight per wave), and the cells above and
// below are kept in registers (a vertical sliding window), so each field is
// fetched from HBM once per sweep.  Left/right neighbours are re-read through
// L1/L2.
//
// The CG iteration kernels (k_cg_pwv, k_cg_urv, k_cgg) are vectorised: a
// thread owns V = 4 (FP32) or 2 (FP64) consecutive columns loaded as one
// 16-byte vector, left/right neighbours come from the adjacent lanes (lanes
// 0 and 63 load the column outside the wave together with the row loads),
// and the rows of a band are loaded in batches that are software-pipelined
// (row_batches): the next batch's loads are issued before the current batch
// is computed, so a wave keeps loads in flight while it computes.
//
// Halo.  The depth-1 reflective halo of kernels.jl:191-210 (with the index
// correction of the intended algorithm) is applied by clamping neighbour
// indices to the interior: a reflected ghost cell holds exactly the value of
// the interior cell it mirrors, so the arithmetic is bit-identical to reading
// an updated halo and no halo exchange is needed inside a solve.
//
// Reductions need no atomics and no fences.  A producing kernel writes one
// partial sum per block (wave64 butterfly, then a fixed-order block sum).
// The consuming kernel -- the next one on the stream, so the kernel boundary
// makes the partials visible -- has every block sum the ~2k partials in the
// same fixed order, so all blocks derive the identical scalar (alpha, beta,
// the stop test) and results are bit-reproducible run to run.  Block 0 also
// records the scalars in the device control block for the host.  A stopped
// solve is expressed by Ctrl::stop_iter: launches for later iterations exit
// at once.
#include <algorithm>
#include <type_traits>
#include <climits>
#include <cmath>
#include <mutex>

#include "tl_internal.hpp"

namespace tl {

Geo make_geo(int nx, int ny, int hd)
{
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.hd = hd;
    g.X = nx + 2 * hd;
    g.Y = ny + 2 * hd;
    g.pitch = ((long)(nx + hd + XCOL) + 63) / 64 * 64;
    g.off = (long)hd * g.pitch + XCOL;
    g.elems = (size_t)g.pitch * (size_t)g.Y + 64;
    g.gx = (nx + TB - 1) / TB;
    // about 2048 blocks on large grids (8 per CU): enough waves in flight,
    // few enough partials for every consumer block to sum them
    const long rows = (long)ny * g.gx;
    int ty = (int)((rows + 2047) / 2048);
    ty = (ty + 7) / 8 * 8;  // a multiple of the row chunk
    if (ty < 8) ty = 8;
    if (ty > 1024) ty = 1024;
    g.ty = ty;
    g.gy = (ny + ty - 1) / ty;
    // vector kernels: the same target block count, 4 (FP32) or 2 (FP64) columns per thread
    // (measured at 8192^2: the FP64 kernels run fastest with about 8192
    // blocks, the FP32 UR with about 2048 and the FP32 PW with about 1024)
    auto vec = [&](int v, long target, int rv, int &gxv, int &gyv, int &tyv) {
        gxv = (nx + v * TBV - 1) / (v * TBV);
        const long r2 = (long)ny * gxv;
        int t2 = (int)((r2 + target - 1) / target);
        t2 = (t2 + rv - 1) / rv * rv;  // a multiple of the kernel's row batch
        // bands of at least 16 rows (each re-reads the row above and below
        // it) unless that leaves fewer than about 1024 blocks
        const long tcap = std::min<long>(16, std::max<long>(rv, r2 / 1024));
        const int tmin = (int)((tcap + rv - 1) / rv * rv);
        if (t2 < tmin) t2 = tmin;
        if (t2 > 1024) t2 = 1024;
        tyv = t2;
        gyv = (ny + t2 - 1) / t2;
    };
    vec(4, TL_VEC_BLOCKS32, TL_RV_UR32, g.gx4, g.gy4, g.ty4);
    int gxp;
    vec(4, TL_VEC_BLOCKS32_PW, TL_RV_PW32, gxp, g.gyp4, g.typ4);  // FP32 PW: same columns, its own bands
    vec(2, TL_VEC_BLOCKS64, TL_RV, g.gx2, g.gy2, g.ty2);
    {
        // fused FP32 kernel: about TL_VEC_BLOCKS_F32 blocks, bands of at
        // least TL_MIN_ROWS_F32 rows (a multiple of its row batch)
        const long r2 = (long)ny * g.gx4;
        int t2 = (int)((r2 + TL_VEC_BLOCKS_F32 - 1) / TL_VEC_BLOCKS_F32);
        t2 = std::max(t2, std::min(TL_MIN_ROWS_F32, ny));
        t2 = (t2 + TL_RV_F32 - 1) / TL_RV_F32 * TL_RV_F32;
        g.tyf = std::min(t2, 1024);
        g.gyf = (ny + g.tyf - 1) / g.tyf;
    }
    {
        // two-step PPCG kernel: about TL_VEC_BLOCKS_PPCG2 blocks (each band
        // re-reads 2 rows below and 2 above it), bands of at least 16 rows
        const long r2 = (long)ny * g.gx2;
        int t2 = (int)((r2 + TL_VEC_BLOCKS_PPCG2 - 1) / TL_VEC_BLOCKS_PPCG2);
        t2 = std::max(t2, std::min(16, ny));
        g.tyq = std::min(t2, 1024);
        g.gyq = (ny + g.tyq - 1) / g.tyq;
    }
    g.nbmax = std::max(std::max(g.gx * g.gy, g.gx4 * g.gyp4), std::max(g.gx4 * g.gy4, g.gx2 * g.gy2));
    g.nbmax = std::max(g.nbmax, g.gx4 * g.gyf);
    return g;
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned block_id() { return blockIdx.y * gridDim.x + blockIdx.x; }

// Partial sums cross kernel boundaries between blocks that run on different
// XCDs, each with its own L2.  They are stored with agent-scope atomic
// stores, so a consumer that reads them with agent-scope atomic loads (the
// single-block finalizers, COH below) sees them whatever the dispatch
// packets' cache fences do; the stencil kernels read them with plain loads
// and rely on the kernel-boundary acquire, as stream order guarantees.
__device__ __forceinline__ void store_partial(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one partial per block: part[a * stride + block] = sum over the block of v[a]
template <int N, int NT = TB>
__device__ __forceinline__ void write_partials(const double (&v)[N], double *__restrict__ part, size_t stride)
{
    __shared__ double sh[N][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        const double s = wave_sum(v[a]);
        if (lane == 0) sh[a][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned b = block_id();
#pragma unroll
        for (int a = 0; a < N; ++a) {
            double s = sh[a][0];
#pragma unroll
            for (int q = 1; q < NT / 64; ++q) s += sh[a][q];
            store_partial(&part[(size_t)a * stride + b], s);
        }
    }
}

// every thread of the block gets the fixed-order sum of nb partials;
// COH: agent-scope atomic loads (coherent across the XCDs' L2s)
template <int N, int NT = TB, bool COH = false>
__device__ __forceinline__ void sum_partials(const double *__restrict__ part, unsigned nb, size_t stride,
                                             double (&tot)[N])
{
    __shared__ double sh[N][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        double s = 0.0;
        for (unsigned q0 = threadIdx.x; q0 < nb; q0 += 8 * NT) {
            double v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const unsigned q = q0 + (unsigned)k * NT;
                const double *pq = part + (size_t)a * stride + q;
                if constexpr (COH)
                    v[k] = q < nb ? __hip_atomic_load(pq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
                else
                    v[k] = q < nb ? *pq : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) s += v[k];
        }
        s = wave_sum(s);
        if (lane == 0) sh[a][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < N; ++a) {
        double s = sh[a][0];
#pragma unroll
        for (int q = 1; q < NT / 64; ++q) s += sh[a][q];
        tot[a] = s;
    }
    __syncthreads();  // sh may be reused by the next call
}

__device__ __forceinline__ unsigned nblocks(const Geo &g) { return (unsigned)(g.gx * g.gy); }

// ------------------------------------------------------------- operator

// smvp, kernels.jl:178-184, intended 5-point form; association as written
template <typename T>
__device__ __forceinline__ T smvp5(T c, T l, T r, T d, T u, T kxl, T kxr, T kyd, T kyu)
{
    T consum = (T)1 + kxr;
    consum = consum + kxl;
    consum = consum + kyu;
    consum = consum + kyd;
    return (consum * c - (kxr * r + kxl * l)) - (kyu * u + kyd * d);
}

// FP32 (inner CG of the mixed-precision solver): the same operator in flux
// form, c + sum k (c - nbr).  The differences of neighbouring values are
// formed before the face coefficients scale them, so the rounding error of
// A p is relative to the differences rather than to (1 + sum k) |p|, which
// for smooth p (the slowest CG modes here, where sum k reaches 1e6) is the
// size of A p itself in the diagonal form.
template <>
__device__ __forceinline__ float smvp5<float>(float c, float l, float r, float d, float u, float kxl, float kxr,
                                              float kyd, float kyu)
{
    return c + ((kxr * (c - r) + kxl * (c - l)) + (kyu * (c - u) + kyd * (c - d)));
}

// One thread owns one column of a tile: rows [y0, y1) of column x.  Rows are
// processed in chunks of R whose loads are all issued before any arithmetic,
// so every wave keeps R rows of loads in flight.  Row indices are clamped
// into [0, ny-1]: the clamped read of row -1 or ny IS the reflective ghost
// value, and rows past the tile end are loaded (valid addresses) but never
// stored.
constexpr int R = 8;

struct Col {
    int x, y0, y1;
    long xl, xr;
    bool active;
};

__device__ __forceinline__ Col column(const Geo &g)
{
    Col c;
    c.x = blockIdx.x * TB + threadIdx.x;
    c.y0 = blockIdx.y * g.ty;
    c.y1 = min(c.y0 + g.ty, g.ny);
    c.active = c.x < g.nx && c.y0 < g.ny;
    c.xl = c.x > 0 ? -1 : 0;
    c.xr = c.x < g.nx - 1 ? 1 : 0;
    return c;
}

// element index of interior cell (x, y)
__device__ __forceinline__ long at(const Geo &g, int x, int y) { return g.off + (long)y * g.pitch + x; }
// same with y clamped into the interior
__device__ __forceinline__ long atc(const Geo &g, int x, int y) { return at(g, x, min(y, g.ny - 1)); }

// stencil inputs of R rows of field a: the row above, left, right, and the
// face coefficients kx(x), kx(x+1), ky(y+1); ky(y) is carried from the row below
template <typename T>
struct Stn {
    T up[R], lf[R], rt[R], kxl[R], kxr[R], kyu[R];
};

template <typename T>
__device__ __forceinline__ void load_stn(Stn<T> &s, const Geo &g, const Col &c, int yb, const T *__restrict__ a,
                                         const T *__restrict__ kx, const T *__restrict__ ky)
{
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int y = yb + q;
        const long ic = atc(g, c.x, y);
        s.up[q] = a[atc(g, c.x, y + 1)];
        s.lf[q] = a[ic + c.xl];
        s.rt[q] = a[ic + c.xr];
        s.kxl[q] = kx[ic];
        s.kxr[q] = kx[ic + 1];
        s.kyu[q] = ky[at(g, c.x, min(y + 1, g.ny))];
    }
}

// vertical window: value at the current row and the row below, ky below
template <typename T>
struct Win {
    T dn, cur, kyd;
    __device__ __forceinline__ void start(const Geo &g, const Col &c, const T *__restrict__ a,
                                          const T *__restrict__ ky)
    {
        cur = a[at(g, c.x, c.y0)];
        dn = a[at(g, c.x, max(c.y0 - 1, 0))];
        kyd = ky[at(g, c.x, c.y0)];
    }
    __device__ __forceinline__ T apply(const Stn<T> &s, int q) const
    {
        return smvp5(cur, s.lf[q], s.rt[q], dn, s.up[q], s.kxl[q], s.kxr[q], kyd, s.kyu[q]);
    }
    __device__ __forceinline__ void advance(const Stn<T> &s, int q)
    {
        dn = cur;
        cur = s.up[q];
        kyd = s.kyu[q];
    }
};

// R values of field a at rows yb..yb+R-1 (clamped)
template <typename T>
__device__ __forceinline__ void load_rows(T (&v)[R], const Geo &g, const Col &c, int yb, const T *__restrict__ a)
{
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = a[atc(g, c.x, yb + q)];
}

// ------------------------------------------------------- scalar updates

// stop test after iteration t with squared residual rr (written by block 0)
__device__ int decide_why(const Ctrl *c, int t, double rr)
{
    const double nr = sqrt(fabs(rr));
    int why = WHY_NONE;
    switch (c->mode) {
    case MODE_CG_REF:
        if (nr < c->eps) why = WHY_CONVERGED;  // CG.jl:24
        break;
    case MODE_CG_REL:
        if (nr <= c->tol_abs) why = WHY_CONVERGED;
        break;
    case MODE_SWITCH: {
        if (fabs(rr) < c->eps) {  // Cheby.jl:57, PPCG.jl:50
            why = WHY_CONVERGED;
        } else {
            const int tt = t + 1;  // the switch test at the top of the next iteration
            const bool sw = c->errorswitch ? (rr < c->epslim && tt > 20) : (tt > c->presteps && rr < 1.0);
            if (sw) why = WHY_SWITCH;
        }
        break;
    }
    case MODE_IR:
        if (nr <= c->tol_abs || nr < c->refresh_abs) why = WHY_REFRESH;
        break;
    case MODE_PPCG:
    case MODE_JACOBI:
        if (fabs(rr) < c->eps) why = WHY_CONVERGED;
        break;
    }
    if (why == WHY_NONE && t >= c->maxiters) why = WHY_MAXITERS;
    if (!isfinite(rr)) why = WHY_BREAKDOWN;
    return why;
}

__device__ void record_stop(Ctrl *c, int t, int why)
{
    if (why == WHY_NONE) return;
    c->why = why;
    if (why == WHY_BREAKDOWN) c->status = TL_E_BREAKDOWN;
    c->stop_iter = t;
}

// Iteration number of a CG iteration kernel.  Stream launches pass t >= 1.
// The launches of a hipGraph batch pass -i for the i-th iteration of the
// batch and take t = Ctrl::tg + i, tg being set by the batch's first node
// (k_graph_base), so one captured graph serves every batch.
__device__ __forceinline__ int iter_of(int targ, const Ctrl *c) { return targ > 0 ? targ : c->tg - targ; }

// first node of a hipGraph batch of k iterations: tg = first iteration of
// this replay, tgn = first iteration of the next one
__global__ __launch_bounds__(64) void k_graph_base(Ctrl *c, int k)
{
    if (threadIdx.x != 0) return;
    const int t0 = c->tgn;
    c->tg = t0;
    c->tgn = t0 + k;
}

// before the first replay of a solve: the first batch starts at iteration t0
__global__ __launch_bounds__(64) void k_graph_seed(Ctrl *c, int t0)
{
    if (threadIdx.x == 0) c->tgn = t0;
}

// -------------------------------------------------------- init kernels

__global__ __launch_bounds__(TB) void k_init_ctrl(Ctrl *c, int mode, int maxiters, double eps, double rel_tol,
                                                  double ir_delta, int presteps, int errorswitch, double epslim)
{
    if (threadIdx.x != 0) return;
    c->rro = c->pw = c->alpha = c->beta = 0.0;
    c->rrs[0] = c->rrs[1] = c->alphas[0] = c->alphas[1] = 0.0;
    c->err = 1e10; /* ERROR_START, kernels.jl:7 */
    c->bb = 0.0;
    c->eps = eps;
    c->rel_tol = rel_tol;
    c->ir_delta = ir_delta;
    c->tol_abs = 0.0;
    c->refresh_abs = 0.0;
    c->epslim = epslim;
    c->iters = 0;
    c->maxiters = maxiters;
    c->mode = mode;
    c->presteps = presteps;
    c->errorswitch = errorswitch;
    c->stop_iter = INT_MAX;
    c->why = WHY_NONE;
    c->status = 0;
}

// switches the stop rule between phases and lets iterations run again
__global__ __launch_bounds__(TB) void k_set_mode(Ctrl *c, int mode, int maxiters)
{
    if (threadIdx.x != 0) return;
    c->mode = mode;
    c->maxiters = maxiters;
    c->stop_iter = INT_MAX;
    c->why = WHY_NONE;
}

// Single-block finalizers: sum n partial arrays (each nb long, stride nb)
// into ctrl->sums[slot ...] and act on them.
enum FinKind : int { FK_SUMS = 0, FK_INIT = 1, FK_BB = 2, FK_CHEBY = 3, FK_JACOBI = 4, FK_IR = 5 };

__global__ __launch_bounds__(TB) void k_fin(Ctrl *__restrict__ c, const double *__restrict__ part, unsigned nb,
                                            unsigned stride, int n, int slot, int kind, int t)
{
    if ((kind == FK_CHEBY || kind == FK_JACOBI) && t > c->stop_iter) return;
    double tot[2] = {0.0, 0.0};
    if (n >= 2) {
        sum_partials<2, TB, true>(part, nb, stride, tot);
    } else {
        double t1[1];
        sum_partials<1, TB, true>(part, nb, stride, t1);
        tot[0] = t1[0];
    }
    double t34[2] = {0.0, 0.0};
    if (n == 4) sum_partials<2, TB, true>(part + 2 * (size_t)stride, nb, stride, t34);
    if (threadIdx.x != 0) return;
    const double v[4] = {tot[0], tot[1], t34[0], t34[1]};
    for (int a = 0; a < n && slot + a < 4; ++a) c->sums[slot + a] = v[a];
    switch (kind) {
    case FK_INIT:  // after CG.init!: rro = sum r^2, bb = sum u0^2
        c->rro = v[0];
        c->rrs[0] = v[0];
        c->err = v[0];
        c->bb = v[1];
        c->tol_abs = c->rel_tol * sqrt(v[1]);
        c->refresh_abs = c->ir_delta * sqrt(v[0]);
        if (v[0] == 0.0) {  // exact solution already: nothing to iterate
            // (tl_solve recomputes the residual of the returned u with other
            // kernels and rejects the claim if it is not zero: a reduction
            // that missed its partials cannot pass as an exact start)
            c->why = WHY_CONVERGED;
            c->stop_iter = 0;
        }
        break;
    case FK_BB:
        c->bb = v[0];
        break;
    case FK_CHEBY:  // calc2norm iteration t of Cheby.mainstep! (Cheby.jl:102-104, 57)
        c->err = v[0];
        c->iters = t;
        if (fabs(v[0]) < c->eps) {
            c->why = WHY_CONVERGED;
            c->stop_iter = t;
        }
        break;
    case FK_JACOBI: {  // Jacobi.jl:15-26: err = sum |u'-u| (+ sum r^2 every 50)
        const double err = n >= 2 ? v[0] + v[1] : v[0];
        c->err = err;
        c->iters = t;
        record_stop(c, t, decide_why(c, t, err));
        break;
    }
    case FK_IR: {  // FP64 refresh of the mixed-precision CG: converged or resume (t = last iteration)
        const double rr = v[0];
        c->rro = rr;
        c->rrs[t & 1] = rr;
        // the x update of iteration t is in u now: iteration t+1 must not apply it again
        c->alphas[t & 1] = 0.0;
        c->err = rr;
        c->refresh_abs = c->ir_delta * sqrt(rr);
        if (!isfinite(rr)) {
            c->why = WHY_BREAKDOWN;
            c->status = TL_E_BREAKDOWN;
            c->stop_iter = c->iters;
        } else if (sqrt(rr) <= c->tol_abs) {
            c->why = WHY_CONVERGED;
            c->stop_iter = c->iters;
        } else if (c->iters >= c->maxiters) {
            c->why = WHY_MAXITERS;
            c->stop_iter = c->iters;
        } else {
            c->why = WHY_NONE;
            c->stop_iter = INT_MAX;
        }
        break;
    }
    default:
        break;
    }
}

// setchunkstate!, chunk.jl:122-151: rasterise the states in order over every
// cell including the halo; then u = energy0*density on halo(chunk, 1).
__global__ __launch_bounds__(TB) void k_set_state(Geo g, const tl_state *__restrict__ st, int ns, double xmin,
                                                  double ymin, double dx, double dy, double *__restrict__ rho,
                                                  double *__restrict__ e0, double *__restrict__ u)
{
    const int k = blockIdx.x * TB + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= g.X || j >= g.Y) return;
    // vertexx = xmin + dx*((1:x+1) - hd - 1), chunk.jl:76-77 (0-based i - hd)
    const double vx0 = xmin + dx * (double)(k - g.hd), vx1 = xmin + dx * (double)(k + 1 - g.hd);
    const double vy0 = ymin + dy * (double)(j - g.hd), vy1 = ymin + dy * (double)(j + 1 - g.hd);
    const double cx = 0.5 * (vx0 + vx1), cy = 0.5 * (vy0 + vy1);
    double d = st[0].density, e = st[0].energy;
    for (int s = 0; s < ns; ++s) {
        bool apply;
        if (st[s].geometry == 0) {
            apply = vx1 >= st[s].xmin && vx0 < st[s].xmax && vy1 >= st[s].ymin && vy0 < st[s].ymax;
        } else if (st[s].geometry == 1) {
            const double ddx = cx - st[s].xmin, ddy = cy - st[s].ymin;
            apply = ddx * ddx + ddy * ddy <= st[s].radius * st[s].radius;
        } else {
            apply = vx0 == st[s].xmin && vy0 == st[s].ymin;
        }
        if (apply) {
            d = st[s].density;
            e = st[s].energy;
        }
    }
    const long i = (long)j * g.pitch + k + (XCOL - g.hd);
    rho[i] = d;
    e0[i] = e;
    if (k >= 1 && k < g.X - 1 && j >= 1 && j < g.Y - 1) u[i] = e * d;
}

// elementwise over the whole X*Y array
__global__ __launch_bounds__(TB) void k_copy_full(Geo g, const double *__restrict__ a, double *__restrict__ b)
{
    const int k = blockIdx.x * TB + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= g.X || j >= g.Y) return;
    const long i = (long)j * g.pitch + k + (XCOL - g.hd);
    b[i] = a[i];
}

// plain 16-byte-per-lane copy of n2 double pairs (bandwidth calibration)
__global__ __launch_bounds__(TB) void k_copy_vec(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n2)
{
    const size_t i = (size_t)blockIdx.x * TB + threadIdx.x;
    if (i < n2) b[i] = a[i];
}

// CG.init! lines 51-68 (CG.jl) / Jacobi.init! lines 38-55 (Jacobi.jl):
// u = energy*density on every cell; p = r = 0 (CG) or u0 = u (Jacobi); the
// face coefficients kx, ky over the reference's range k, j in hd+1 : X-1.
__global__ __launch_bounds__(TB) void k_init_coef(Geo g, const double *__restrict__ rho,
                                                  const double *__restrict__ e, double *__restrict__ u,
                                                  double *__restrict__ p, double *__restrict__ r,
                                                  double *__restrict__ u0, double *__restrict__ kx,
                                                  double *__restrict__ ky, double rx, double ry, int recip,
                                                  int what)
{
    const int k = blockIdx.x * TB + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= g.X || j >= g.Y) return;
    const long i = (long)j * g.pitch + k + (XCOL - g.hd);
    if (what != INIT_COEF_ONLY) {
        const double uv = e[i] * rho[i];
        u[i] = uv;
        if (what == INIT_JACOBI) {
            u0[i] = uv;
        } else {
            p[i] = 0.0;
            r[i] = 0.0;
        }
    }
    if (k >= g.hd && k <= g.X - 2 && j >= g.hd && j <= g.Y - 2) {
        const double wc = recip ? 1.0 / rho[i] : rho[i];
        const double wl = recip ? 1.0 / rho[i - 1] : rho[i - 1];
        const double wd = recip ? 1.0 / rho[i - g.pitch] : rho[i - g.pitch];
        kx[i] = rx * (wl + wc) / (2.0 * wl * wc);
        ky[i] = ry * (wd + wc) / (2.0 * wd * wc);
    }
}

// ------------------------------------------------------------ CG family

// CG.init! lines 70-78 + copyu! (kernels.jl:217-220), with an optional
// initial guess: b = u (energy*density); x = guess or b;
// w = A x; r = b - w; p = r; u0 = b; u = x.  Partials of sum r^2 go to
// pr0 (the "previous rr" of iteration 1) and pm[0]; sum b^2 to pm[1].
__global__ __launch_bounds__(TB) void k_cg_init_resid(Geo g, double *__restrict__ u,
                                                      const double *__restrict__ guess, double *__restrict__ u0,
                                                      double *__restrict__ p, double *__restrict__ r,
                                                      double *__restrict__ w, const double *__restrict__ kx,
                                                      const double *__restrict__ ky, double *__restrict__ pr0,
                                                      double *__restrict__ pm)
{
    const Col c = column(g);
    double acc[2] = {0.0, 0.0};
    const double *xs = guess ? guess : u;
    if (c.active) {
        Win<double> win;
        win.start(g, c, xs, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<double> st;
            double b[R];
            load_stn(st, g, c, yb, xs, kx, ky);
            load_rows(b, g, c, yb, u);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double wv = win.apply(st, q);
                const double rv = b[q] - wv;
                if (yb + q < c.y1) {
                    const long i = at(g, c.x, yb + q);
                    w[i] = wv;
                    r[i] = rv;
                    p[i] = rv;
                    u0[i] = b[q];
                    if (guess) u[i] = win.cur;
                    acc[0] += rv * rv;
                    acc[1] += b[q] * b[q];
                }
                win.advance(st, q);
            }
        }
    }
    write_partials<2>(acc, pm, g.nbmax);
    if (threadIdx.x == 0) pr0[block_id()] = pm[block_id()];
}

// CG.w!, CG.jl:82-90: w = A p; partial sums of p*w
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_w(Geo g, int t, const T *__restrict__ p, const T *__restrict__ kx,
                                             const T *__restrict__ ky, T *__restrict__ w,
                                             const Ctrl *__restrict__ ctrl, double *__restrict__ pw)
{
    if (t > ctrl->stop_iter) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        Win<T> win;
        win.start(g, c, p, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<T> st;
            load_stn(st, g, c, yb, p, kx, ky);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const T wv = win.apply(st, q);
                if (yb + q < c.y1) {
                    w[at(g, c.x, yb + q)] = wv;
                    acc[0] += (double)wv * (double)win.cur;
                }
                win.advance(st, q);
            }
        }
    }
    write_partials<1>(acc, pw, g.nbmax);
}

// alpha = rro / pw (CG.jl:35), derived identically by every block
template <int NT = TB>
__device__ __forceinline__ bool cg_alpha(const Geo &g, int t, Ctrl *__restrict__ ctrl,
                                         const double *__restrict__ pw, unsigned nbw,
                                         const double *__restrict__ prprev, unsigned nbp,
                                         double *__restrict__ hist_a, double *alpha)
{
    double a1[1], a2[1];
    sum_partials<1, NT>(pw, nbw, g.nbmax, a1);
    sum_partials<1, NT>(prprev, nbp, g.nbmax, a2);
    const double pwv = a1[0], rro = a2[0];
    const double a = rro / pwv;
    const bool ok = pwv > 0.0 && isfinite(a);
    if (block_id() == 0 && threadIdx.x == 0) {
        ctrl->pw = pwv;
        ctrl->alpha = a;
        if (hist_a) hist_a[t - 1] = a;
        if (!ok) {  // non-positive curvature: stop before this iteration's update
            ctrl->status = TL_E_BREAKDOWN;
            ctrl->why = WHY_BREAKDOWN;
            ctrl->stop_iter = t - 1;
        }
    }
    *alpha = a;
    return ok;
}

// alpha = rr(t-1) / pw(t) for the vectorised kernels: pw summed from the
// partials, rr(t-1) from the slot the previous reduction left in Ctrl
template <int NT>
__device__ __forceinline__ bool cg_alpha_slot(const Geo &g, int t, Ctrl *__restrict__ ctrl,
                                              const double *__restrict__ pw, unsigned nbw,
                                              double *__restrict__ hist_a, double *alpha)
{
    double a1[1];
    sum_partials<1, NT>(pw, nbw, g.nbmax, a1);
    const double pwv = a1[0], rro = ctrl->rrs[(t - 1) & 1];
    const double a = rro / pwv;
    const bool ok = pwv > 0.0 && isfinite(a);
    if (block_id() == 0 && threadIdx.x == 0) {
        ctrl->pw = pwv;
        ctrl->alpha = a;
        ctrl->alphas[t & 1] = a;
        if (hist_a) hist_a[t - 1] = a;
        if (!ok) {  // non-positive curvature: stop before this iteration's update
            ctrl->status = TL_E_BREAKDOWN;
            ctrl->why = WHY_BREAKDOWN;
            ctrl->stop_iter = t - 1;
        }
    }
    *alpha = a;
    return ok;
}

// CG.ur!, CG.jl:93-98: u += alpha p; r -= alpha w; partial sums of r^2
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_ur(Geo g, int t, T *__restrict__ u, const T *__restrict__ p,
                                              T *__restrict__ r, const T *__restrict__ w, Ctrl *__restrict__ ctrl,
                                              const double *__restrict__ pw, unsigned nbw,
                                              const double *__restrict__ prprev, unsigned nbp,
                                              double *__restrict__ prcur, double *__restrict__ hist_a)
{
    if (t > ctrl->stop_iter) return;
    double ad;
    if (!cg_alpha(g, t, ctrl, pw, nbw, prprev, nbp, hist_a, &ad)) return;
    const T alpha = (T)ad;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        for (int yb = c.y0; yb < c.y1; yb += R) {
            T uv[R], pv[R], rv[R], wv[R];
            load_rows(uv, g, c, yb, u);
            load_rows(pv, g, c, yb, p);
            load_rows(rv, g, c, yb, r);
            load_rows(wv, g, c, yb, w);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (yb + q < c.y1) {
                    const long i = at(g, c.x, yb + q);
                    u[i] = uv[q] + alpha * pv[q];
                    const T rn = rv[q] - alpha * wv[q];
                    r[i] = rn;
                    acc[0] += (double)rn * (double)rn;
                }
            }
        }
    }
    write_partials<1>(acc, prcur, g.nbmax);
}

// beta = rrn / rro (CG.jl:39) and the stop test, derived identically by
// every block; block 0 records them
__device__ __forceinline__ double cg_beta(const Geo &g, int t, Ctrl *__restrict__ ctrl,
                                          const double *__restrict__ prprev, unsigned nbp,
                                          const double *__restrict__ prcur, unsigned nbc,
                                          double *__restrict__ hist_b)
{
    double a1[1], a2[1];
    sum_partials<1>(prcur, nbc, g.nbmax, a1);
    sum_partials<1>(prprev, nbp, g.nbmax, a2);
    const double rrn = a1[0], rro = a2[0];
    const double beta = rrn / rro;
    if (block_id() == 0 && threadIdx.x == 0) {
        ctrl->beta = beta;
        ctrl->rro = rrn;
        ctrl->err = rrn;
        ctrl->iters = t;
        if (hist_b) hist_b[t - 1] = beta;
        record_stop(ctrl, t, decide_why(ctrl, t, rrn));
    }
    return beta;
}

// CG.p!, CG.jl:101-104: p = beta p + r (runs in the iteration that stops)
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_p(Geo g, int t, T *__restrict__ p, const T *__restrict__ r,
                                             Ctrl *__restrict__ ctrl, const double *__restrict__ prprev, unsigned nbp,
                                             const double *__restrict__ prcur, unsigned nbc,
                                             double *__restrict__ hist_b)
{
    if (t > ctrl->stop_iter) return;
    const T beta = (T)cg_beta(g, t, ctrl, prprev, nbp, prcur, nbc, hist_b);
    const Col c = column(g);
    if (!c.active) return;
    for (int yb = c.y0; yb < c.y1; yb += R) {
        T pv[R], rv[R];
        load_rows(pv, g, c, yb, p);
        load_rows(rv, g, c, yb, r);
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (yb + q < c.y1) p[at(g, c.x, yb + q)] = beta * pv[q] + rv[q];
    }
}


// ------------------------------------------- vectorised CG (fused p update)
//
// A thread owns V consecutive columns (16-byte loads: V = 4 for FP32, 2 for
// FP64); a block of TB threads owns V*TB columns and ty rows, processed in
// chunks of RV rows whose loads are issued together.  The left/right
// neighbours of a thread's columns come from the adjacent lanes through wave
// shuffles (lanes 0 and 63 read them from memory).  All lanes run every row
// so the shuffles see the whole wave; lanes past the grid's right edge load a
// valid clamped vector and store nothing.  When nx is not a multiple of V the
// last vector of a row reads into the right halo/padding (inside the
// allocation) and stores only its interior columns.

// NT: non-temporal (streaming) access, the default for the stencil kernels'
// field sweeps (TL_NT_LOAD / TL_NT_STORE)
template <typename T, int V, bool NT = (TL_NT_LOAD != 0)>
__device__ __forceinline__ void vload(T (&d)[V], const T *__restrict__ p)
{
    typedef T vec __attribute__((ext_vector_type(V)));
    vec v;
    if constexpr (NT)
        v = __builtin_nontemporal_load(reinterpret_cast<const vec *>(p));
    else
        v = *reinterpret_cast<const vec *>(p);
#pragma unroll
    for (int j = 0; j < V; ++j) d[j] = v[j];
}

template <typename T, int V, bool NT = (TL_NT_STORE != 0)>
__device__ __forceinline__ void vstore(T *__restrict__ p, const T (&s)[V])
{
    typedef T vec __attribute__((ext_vector_type(V)));
    vec v;
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = s[j];
    if constexpr (NT)
        __builtin_nontemporal_store(v, reinterpret_cast<vec *>(p));
    else
        *reinterpret_cast<vec *>(p) = v;
}

// store of the first n (< V: the grid's last, partial vector) or all V elements
template <typename T, int V, bool NT = (TL_NT_STORE != 0)>
__device__ __forceinline__ void vstore_n(T *__restrict__ p, const T (&s)[V], int n)
{
    if (n >= V) {
        vstore<T, V, NT>(p, s);
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j)
            if (j < n) p[j] = s[j];
    }
}

template <int V>
struct VCol {
    int x0;        // first column owned (may be >= nx: inactive)
    int xl;        // first column loaded (clamped into the grid)
    int y0, y1;
    bool active;
    int lane;
};

// Tile of this block.  The vector kernels launch a 1-D grid of gx*gy blocks
// that the hardware deals round-robin to the 8 XCDs (block b to XCD b % 8),
// each XCD with its own L2.  Row-major tiles would put every horizontally
// adjacent tile on another XCD, so the neighbour columns a block reads at its
// edges come from another L2.  TL_XCD_MAP 2 (the default) gives the 8 bands
// of a group to the 8 XCDs and a band's tiles to consecutive slots of one
// XCD; the blocks in flight still cover a contiguous range of rows.
// TL_XCD_MAP 1 gives each XCD a contiguous slab of bands (slower).
template <int V>
__device__ __forceinline__ void vtile(const Geo &g, int &bx, int &by)
{
    const int gxv = V == 4 ? g.gx4 : g.gx2;
    const unsigned b = blockIdx.x;
    unsigned tile = b;
#if TL_XCD_MAP == 1
    const unsigned nb = gridDim.x;
    if (nb % 8 == 0) tile = (b % 8) * (nb / 8) + b / 8;
#elif TL_XCD_MAP == 2
    // groups of 8 bands: band 8q + x of group q on XCD x = b % 8, its tiles
    // in consecutive slots, so horizontally adjacent tiles share an L2
    const unsigned nb = gridDim.x, gyv = nb / (unsigned)gxv;
    if (gyv % 8 == 0) {
        const unsigned x = b % 8, s = b / 8;
        tile = ((s / (unsigned)gxv) * 8 + x) * (unsigned)gxv + s % (unsigned)gxv;
    }
#endif
    bx = (int)(tile % (unsigned)gxv);
    by = (int)(tile / (unsigned)gxv);
}

template <int V>
__device__ __forceinline__ VCol<V> vcolumn(const Geo &g, int ty)
{
    VCol<V> c;
    int bx, by;
    vtile<V>(g, bx, by);
    c.x0 = (bx * TBV + threadIdx.x) * V;
    c.xl = min(c.x0, ((g.nx - 1) / V) * V);
    c.y0 = by * ty;
    c.y1 = min(c.y0 + ty, g.ny);
    c.active = c.x0 < g.nx;
    c.lane = threadIdx.x & 63;
    return c;
}

// v of lane l-1 (up1) / l+1 (down1) in a wave of 64, by LDS shuffles or (DPP)
// by DPP wave shifts, which move data between VALU lanes without the LDS
// crossbar (lane 0 / 63 then keep their own value; the callers replace it).
// Measured at 8192^2: DPP makes the FP32 PW kernel 4% faster and the others
// no faster, so only that kernel uses it.
__device__ __forceinline__ int dpp_up1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_down1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
template <bool DPP>
__device__ __forceinline__ float lane_up1(float v)
{
    if (DPP) {
        return __int_as_float(dpp_up1(__float_as_int(v)));
    }
    return __shfl_up(v, 1, 64);
}
template <bool DPP>
__device__ __forceinline__ float lane_down1(float v)
{
    if (DPP) {
        return __int_as_float(dpp_down1(__float_as_int(v)));
    }
    return __shfl_down(v, 1, 64);
}
template <bool DPP>
__device__ __forceinline__ double lane_up1(double v)
{
    if (DPP) {
        const long long b = __double_as_longlong(v);
        const int lo = dpp_up1((int)(b & 0xffffffffLL)), hi = dpp_up1((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
    return __shfl_up(v, 1, 64);
}
template <bool DPP>
__device__ __forceinline__ double lane_down1(double v)
{
    if (DPP) {
        const long long b = __double_as_longlong(v);
        const int lo = dpp_down1((int)(b & 0xffffffffLL)), hi = dpp_down1((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
    return __shfl_down(v, 1, 64);
}

// values of the columns left and right of the thread's V columns, given the
// thread's own row values c[] (element index i of column xl)
template <typename T, int V, bool DPP = false, class F>
__device__ __forceinline__ void vneigh(const Geo &g, const VCol<V> &cl, const T (&c)[V], F &&at_col, T &lv,
                                       T &rv)
{
    T l = lane_up1<DPP>(c[V - 1]);
    T r = lane_down1<DPP>(c[0]);
    if (cl.lane == 0) l = cl.xl > 0 ? at_col(cl.xl - 1) : c[0];
    if (cl.lane == 63) r = cl.xl + V < g.nx ? at_col(cl.xl + V) : c[V - 1];
    if (cl.xl == 0) l = c[0];
    lv = l;
    rv = r;
}

// 5-point operator on element j of a V-vector row: left/right from the
// vector or the neighbour values, clamped at the grid edges
template <typename T, int V>
__device__ __forceinline__ T vsmvp(const Geo &g, int xl, int j, const T (&c)[V], T lv, T rv, T dn, T up, T kxl,
                                   T kxr, T kyd, T kyu)
{
    const T l = j > 0 ? c[j - 1] : lv;
    T r = j < V - 1 ? c[j + 1] : rv;
    if (xl + j + 1 >= g.nx) r = c[j];
    return smvp5(c[j], l, r, dn, up, kxl, kxr, kyd, kyu);
}

// Rows y0 .. y1-1 in batches of RB: load(batch, first row) issues a batch's
// loads, compute(batch, first row) uses them.  PIPE: software pipeline --
// the next batch's loads are issued before the current batch is computed,
// into a second register set, so a wave keeps loads in flight.
template <bool PIPE, int RB, class Batch, class Load, class Compute>
__device__ __forceinline__ void row_batches(int y0, int y1, Load &&load, Compute &&compute)
{
    if constexpr (PIPE) {
        Batch b0, b1;
        load(b0, y0);
        for (int yb = y0; yb < y1; yb += 2 * RB) {
            const bool more = yb + RB < y1;
            if (more) load(b1, yb + RB);
            compute(b0, yb);
            if (!more) break;
            if (yb + 2 * RB < y1) load(b0, yb + 2 * RB);
            compute(b1, yb + RB);
        }
    } else {
        for (int yb = y0; yb < y1; yb += RB) {
            Batch b0;
            load(b0, yb);
            compute(b0, yb);
        }
    }
}

// Iteration t of the fused CG: for t >= 2, first the end of iteration t-1 --
// beta = rr(t-1)/rr(t-2) and the stop test (CG.jl:39-42, 24) -- then
// p = r + beta p (CG.p!, written to pout; p is ping-ponged because every
// thread reads its neighbours' old p), then w = A p and the partial sums of
// p.w (CG.w!).  For t = 1, p comes from CG.init!.
template <typename T, int V, int RB, bool PIPE>
__global__ __launch_bounds__(TBV) void k_cg_pwv(Geo g, int ty, int t, const T *__restrict__ pin, T *__restrict__ pout,
                                               const T *__restrict__ r, const T *__restrict__ kx,
                                               const T *__restrict__ ky, T *__restrict__ w, Ctrl *__restrict__ ctrl,
                                               double *__restrict__ pw, const double *__restrict__ prcur, unsigned nbc,
                                               double *__restrict__ hist_b)
{
    t = iter_of(t, ctrl);
    if (t > ctrl->stop_iter) return;
    T beta = (T)0;
    const bool fuse = t >= 2;
    if (fuse) {
        // rr(t-1) from the partials of UR(t-1); rr(t-2) as this kernel's
        // previous launch summed it
        double a1[1];
        sum_partials<1, TBV>(prcur, nbc, g.nbmax, a1);
        const double rrn = a1[0], rro = ctrl->rrs[(t - 2) & 1];
        const double b = rrn / rro;
        const int why = decide_why(ctrl, t - 1, rrn);
        if (block_id() == 0 && threadIdx.x == 0) {
            ctrl->rrs[(t - 1) & 1] = rrn;
            ctrl->beta = b;
            ctrl->rro = rrn;
            ctrl->err = rrn;
            ctrl->iters = t - 1;
            if (hist_b) hist_b[t - 2] = b;
            record_stop(ctrl, t - 1, why);
        }
        if (why != WHY_NONE) return;  // every block derives the same decision
        beta = (T)b;
    }
    const VCol<V> cl = vcolumn<V>(g, ty);
    const T *pp = fuse ? pin : pout;
#if !TL_EDGE_PF
    // p after the update at column x of row element index i
    auto pnew = [&](long i) -> T { return fuse ? r[i] + beta * pp[i] : pp[i]; };
#endif
    double acc[1] = {0.0};
    if (cl.y0 < g.ny) {
        T dn[V], cur[V], kyd[V];
        {
            const long i0 = at(g, cl.xl, cl.y0), id = at(g, cl.xl, max(cl.y0 - 1, 0));
            T a[V], b2[V];
            vload<T, V>(a, pp + i0);
            if (fuse) {
                vload<T, V>(b2, r + i0);
#pragma unroll
                for (int j = 0; j < V; ++j) cur[j] = b2[j] + beta * a[j];
                vload<T, V>(a, pp + id);
                vload<T, V>(b2, r + id);
#pragma unroll
                for (int j = 0; j < V; ++j) dn[j] = b2[j] + beta * a[j];
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) cur[j] = a[j];
                vload<T, V>(dn, pp + id);
            }
            vload<T, V>(kyd, ky + i0);
        }
#if TL_EDGE_PF
        // lanes 0 / 63 of a wave read the column left / right of the wave's
        // columns (and lane 63 the face coefficient right of it) together
        // with the row loads below, not when the row is used
        const bool eneed = (cl.lane == 0 && cl.xl > 0) || (cl.lane == 63 && cl.xl + V < g.nx);
        const long eoff = cl.lane == 0 ? -1 : V;
        T ecur = (T)0;
        if (eneed) {
            const long ie = at(g, cl.xl, cl.y0) + eoff;
            ecur = fuse ? r[ie] + beta * pp[ie] : pp[ie];
        }
#endif
        // the loads of RB rows, issued together
        struct Batch {
            T up[RB][V], kxv[RB][V], kyu[RB][V], rup[RB][V];
#if TL_EDGE_PF
            T eup_p[RB], eup_r[RB], ekx[RB];
#endif
        };
        auto load_batch = [&](Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long ic = atc(g, cl.xl, y), iu = atc(g, cl.xl, y + 1);
                vload<T, V>(B.up[q], pp + iu);
                if (fuse) vload<T, V>(B.rup[q], r + iu);
                vload<T, V>(B.kxv[q], kx + ic);
                vload<T, V>(B.kyu[q], ky + at(g, cl.xl, min(y + 1, g.ny)));
#if TL_EDGE_PF
                B.eup_p[q] = B.eup_r[q] = B.ekx[q] = (T)0;
                if (eneed) {
                    B.eup_p[q] = pp[iu + eoff];
                    if (fuse) B.eup_r[q] = r[iu + eoff];
                }
                if (cl.lane == 63) B.ekx[q] = kx[ic + V];
#endif
            }
        };
        auto compute_batch = [&](const Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long ic = atc(g, cl.xl, y);
                T u2[V];
#pragma unroll
                for (int j = 0; j < V; ++j) u2[j] = fuse ? B.rup[q][j] + beta * B.up[q][j] : B.up[q][j];
                // clamp the row above at the top edge (reflection: the row itself)
                if (y + 1 >= g.ny) {
#pragma unroll
                    for (int j = 0; j < V; ++j) u2[j] = cur[j];
                }
                T lv, rv;
#if TL_EDGE_PF
                auto col_val = [&](int) -> T { return ecur; };
#else
                auto col_val = [&](int x) -> T { return pnew(ic + (x - cl.xl)); };
#endif
                vneigh<T, V>(g, cl, cur, col_val, lv, rv);
                // kx of the face right of the last column: next lane's first kx
                T kxn = lane_down1<false>(B.kxv[q][0]);
#if TL_EDGE_PF
                if (cl.lane == 63) kxn = B.ekx[q];
#else
                if (cl.lane == 63) kxn = kx[ic + V];
#endif
                T wv[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxr = j < V - 1 ? B.kxv[q][j + 1] : kxn;
                    wv[j] = vsmvp<T, V>(g, cl.xl, j, cur, lv, rv, dn[j], u2[j], B.kxv[q][j], kxr, kyd[j],
                                        B.kyu[q][j]);
                }
                if (y < cl.y1 && cl.active) {
                    vstore_n<T, V>(w + ic, wv, g.nx - cl.x0);
                    if (fuse) vstore_n<T, V>(pout + ic, cur, g.nx - cl.x0);
#pragma unroll
                    for (int j = 0; j < V; ++j)
                        if (cl.x0 + j < g.nx) acc[0] += (double)wv[j] * (double)cur[j];
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    dn[j] = cur[j];
                    cur[j] = u2[j];
                    kyd[j] = B.kyu[q][j];
                }
#if TL_EDGE_PF
                // the row above becomes the current row (at the top edge the
                // clamped load read the current row itself)
                ecur = fuse ? B.eup_r[q] + beta * B.eup_p[q] : B.eup_p[q];
#endif
            }
        };
        row_batches<PIPE, RB, Batch>(cl.y0, cl.y1, load_batch, compute_batch);
    }
    write_partials<1, TBV>(acc, pw, g.nbmax);
}

// CG.ur!, CG.jl:93-98 on V-vectors: alpha = rro/pw from the partials, then
// u += alpha p; r -= alpha w; partial sums of r^2
template <typename T, int V, int RB, bool PIPE>
__global__ __launch_bounds__(TBV) void k_cg_urv(Geo g, int ty, int t, T *__restrict__ u, const T *__restrict__ p,
                                               T *__restrict__ r, const T *__restrict__ w, Ctrl *__restrict__ ctrl,
                                               const double *__restrict__ pw, unsigned nbw,
                                               double *__restrict__ prcur, double *__restrict__ hist_a)
{
    t = iter_of(t, ctrl);
    if (t > ctrl->stop_iter) return;
    double ad;
    if (!cg_alpha_slot<TBV>(g, t, ctrl, pw, nbw, hist_a, &ad)) return;
    const T alpha = (T)ad;
    const VCol<V> cl = vcolumn<V>(g, ty);
    double acc[1] = {0.0};
    if (cl.active && cl.y0 < g.ny) {
        struct Batch {
            T uv[RB][V], pv[RB][V], rv[RB][V], wv[RB][V];
        };
        auto load_batch = [&](Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const long i = atc(g, cl.xl, yb + q);
                vload<T, V>(B.uv[q], u + i);
                vload<T, V>(B.pv[q], p + i);
                vload<T, V>(B.rv[q], r + i);
                vload<T, V>(B.wv[q], w + i);
            }
        };
        auto compute_batch = [&](const Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                if (yb + q < cl.y1) {
                    const long i = at(g, cl.xl, yb + q);
                    T un[V], rn[V];
#pragma unroll
                    for (int j = 0; j < V; ++j) {
                        un[j] = B.uv[q][j] + alpha * B.pv[q][j];
                        rn[j] = B.rv[q][j] - alpha * B.wv[q][j];
                        if (cl.x0 + j < g.nx) acc[0] += (double)rn[j] * (double)rn[j];
                    }
                    vstore_n<T, V>(u + i, un, g.nx - cl.x0);
                    vstore_n<T, V>(r + i, rn, g.nx - cl.x0);
                }
            }
        };
        row_batches<PIPE, RB, Batch>(cl.y0, cl.y1, load_batch, compute_batch);
    }
    write_partials<1, TBV>(acc, prcur, g.nbmax);
}

// Cell coefficients.  A coefficient is held in a 32-bit register as loaded
// (16-bit storage: zero-extended) and converted to FP32 only where it is
// used, so that no instruction between the load and the use waits for it.
// FMT (GcFmt) selects half or bfloat16 for 16-bit storage; both convert to
// FP32 exactly.
__device__ __forceinline__ unsigned gc_raw(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned gc_raw(unsigned short h) { return (unsigned)h; }

template <int FMT>
__device__ __forceinline__ float gc_half(unsigned b)  // one 16-bit value in the low bits of b
{
    if constexpr (FMT == GC_F16) {
        return (float)__builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu));
    } else {
        return __uint_as_float(b << 16);
    }
}

template <int FMT>
__device__ __forceinline__ float gc_from_raw(unsigned b)
{
#if TL_GC16
    return gc_half<FMT>(b);
#else
    return __uint_as_float(b);
#endif
}

// V coefficients as loaded (16-bit: two per 32-bit word; the vector is
// aligned like the FP32 vector at the same element index), unpacked where used
constexpr int GCW = TL_GC16 ? 2 : 1;
template <int V>
struct GcRaw {
    unsigned w[V / GCW];
};

template <int V>
__device__ __forceinline__ void gc_load_raw(GcRaw<V> &d, const tl_gc_t *__restrict__ p)
{
    typedef unsigned uv __attribute__((ext_vector_type(V / GCW)));
    const uv v = *reinterpret_cast<const uv *>(p);
#pragma unroll
    for (int k = 0; k < V / GCW; ++k) d.w[k] = v[k];
}

template <int V, int FMT>
__device__ __forceinline__ void gc_unpack(float (&d)[V], const GcRaw<V> &s)
{
#if TL_GC16
#pragma unroll
    for (int k = 0; k < V / 2; ++k) {
        if constexpr (FMT == GC_F16) {
            d[2 * k] = gc_half<FMT>(s.w[k]);
            d[2 * k + 1] = gc_half<FMT>(s.w[k] >> 16);
        } else {
            d[2 * k] = __uint_as_float(s.w[k] << 16);
            d[2 * k + 1] = __uint_as_float(s.w[k] & 0xffff0000u);
        }
    }
#else
#pragma unroll
    for (int k = 0; k < V; ++k) d[k] = __uint_as_float(s.w[k]);
#endif
}

template <int V, int FMT>
__device__ __forceinline__ void gc_load(float (&d)[V], const tl_gc_t *__restrict__ p)
{
    GcRaw<V> w;
    gc_load_raw<V>(w, p);
    gc_unpack<V, FMT>(d, w);
}

// ------------------------------- FP32 inner CG on a cell-coefficient array
//
// The inner CG of the mixed-precision solver builds the operator from one
// array per cell, gc = 1/w (w = density or 1/density, CG.jl:56-58), stored
// as bfloat16 (tl_gc_t), instead of the two FP64 face arrays:
// kx(x) = hx (gc(x-1) + gc(x)) with hx = rx/2 equals
// rx (w(x-1) + w(x)) / (2 w(x-1) w(x)) of CG.jl:61-68 up to the rounding of
// gc (ky likewise with hy = ry/2), computed in FP32.  At the grid edge the
// clamped neighbour supplies what the reflective halo holds
// (kernels.jl:191-210).  w = A p is never stored: UR recomputes it from p.
// Bytes per cell: PW reads p, r (4 each), gc (2) and writes p' (4): 14;
// UR 14 or 26 (below).
//
// UR = false (PW, iteration t): for t >= 2 first the end of iteration t-1 --
//   beta = rr(t-1)/rr(t-2), the stop test (CG.jl:39-42, 24) -- then
//   p = r + beta p (CG.p!, written to pout; p is ping-ponged because every
//   thread reads its neighbours' old p); partial sums of p.(A p) (CG.w!).
// UR = true (iteration t): alpha = rr(t-1)/pw(t), then r -= alpha A p and
//   partial sums of r^2 (CG.ur!, CG.jl:93-98); the correction x += alpha p
//   is applied for two iterations at once, at even t, which saves x's read
//   and write at odd t: UR moves 26 bytes per cell at even t, 14 at odd t.
template <int V, bool UR, int RB, bool PIPE, int FMT>
__global__ __launch_bounds__(TBV) TL_CGG_ATTR void k_cgg(Geo g, int ty, int t, const float *__restrict__ pin,
                                            float *__restrict__ pout, float *__restrict__ r, float *__restrict__ x,
                                            const tl_gc_t *__restrict__ gc, float hx, float hy,
                                            Ctrl *__restrict__ ctrl, const double *__restrict__ pw_in, unsigned nbw,
                                            const double *__restrict__ prcur, unsigned nbc,
                                            double *__restrict__ part_out, const float *__restrict__ pprev)
{
    t = iter_of(t, ctrl);
    if (t > ctrl->stop_iter) return;
    float beta = 0.0f, alpha = 0.0f;
    const bool fuse = !UR && t >= 2;
    // UR at even t also applies the x update of iteration t-1, whose p is
    // still in the other ping-pong buffer: x += alpha(t-1) p(t-1) + alpha(t) p(t)
    const bool xupd = UR && (t & 1) == 0;
    float alpha_prev = 0.0f;
    if (UR) {
        double ad;
        if (!cg_alpha_slot<TBV>(g, t, ctrl, pw_in, nbw, nullptr, &ad)) return;
        alpha = (float)ad;
        if (xupd) alpha_prev = (float)ctrl->alphas[(t - 1) & 1];
    } else if (fuse) {
        // rr(t-1) from the partials of UR(t-1) (or of the FP64 refresh);
        // rr(t-2) as this kernel's previous launch summed it
        double a1[1];
        sum_partials<1, TBV>(prcur, nbc, g.nbmax, a1);
        const double rrn = a1[0], rro = ctrl->rrs[(t - 2) & 1];
        const double b = rrn / rro;
        const int why = decide_why(ctrl, t - 1, rrn);
        if (block_id() == 0 && threadIdx.x == 0) {
            ctrl->rrs[(t - 1) & 1] = rrn;
            ctrl->beta = b;
            ctrl->rro = rrn;
            ctrl->err = rrn;
            ctrl->iters = t - 1;
            record_stop(ctrl, t - 1, why);
        }
        if (why != WHY_NONE) return;  // every block derives the same decision
        beta = (float)b;
    }
    const VCol<V> cl = vcolumn<V>(g, ty);
    const float *pp = fuse ? pin : pout;
#if !TL_EDGE_PF
    // p (after this launch's update) at element i
    auto pnew = [&](long i) -> float { return fuse ? r[i] + beta * pp[i] : pp[i]; };
    auto gcol = [&](long i) -> float { return gc_from_raw<FMT>(gc_raw(gc[i])); };
#endif
    double acc[1] = {0.0};
    if (cl.y0 < g.ny) {
        float dn[V], cur[V], gdn[V], gcur[V];
        {
            const long i0 = at(g, cl.xl, cl.y0), id = at(g, cl.xl, max(cl.y0 - 1, 0));
            float a[V], b2[V];
            vload<float, V>(a, pp + i0);
            gc_load<V, FMT>(gcur, gc + i0);
            gc_load<V, FMT>(gdn, gc + id);
            if (fuse) {
                vload<float, V>(b2, r + i0);
#pragma unroll
                for (int j = 0; j < V; ++j) cur[j] = b2[j] + beta * a[j];
                vload<float, V>(a, pp + id);
                vload<float, V>(b2, r + id);
#pragma unroll
                for (int j = 0; j < V; ++j) dn[j] = b2[j] + beta * a[j];
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) cur[j] = a[j];
                vload<float, V>(dn, pp + id);
            }
        }
#if TL_EDGE_PF
        // lanes 0 / 63 of a wave read the column left / right of the wave's
        // columns together with the row loads below, not when the row is used
        const bool eneed = (cl.lane == 0 && cl.xl > 0) || (cl.lane == 63 && cl.xl + V < g.nx);
        const long eoff = cl.lane == 0 ? -1 : V;
        float ecur_p = 0.0f, ecur_g = 0.0f;
        if (eneed) {
            const long ie = at(g, cl.xl, cl.y0) + eoff;
            ecur_p = fuse ? r[ie] + beta * pp[ie] : pp[ie];
            ecur_g = gc_from_raw<FMT>(gc_raw(gc[ie]));
        }
#endif
        // the loads of RB rows, issued together
        struct Batch {
            float up[RB][V], rup[RB][V], xv[RB][V], rv[RB][V], pv[RB][V];
            GcRaw<V> gup[RB];
#if TL_EDGE_PF
            float eup_p[RB], eup_r[RB];
            // a full register per value: bf16 halves packed into one register
            // would make the compiler wait for each load before the next row's
            unsigned eup_g[RB];
#endif
        };
        auto load_batch = [&](Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long iu = atc(g, cl.xl, y + 1);
                vload<float, V>(B.up[q], pp + iu);
                if (fuse) vload<float, V>(B.rup[q], r + iu);
                gc_load_raw<V>(B.gup[q], gc + iu);
#if TL_EDGE_PF
                B.eup_p[q] = B.eup_r[q] = 0.0f;
                B.eup_g[q] = 0;
                if (eneed) {
                    B.eup_p[q] = pp[iu + eoff];
                    if (fuse) B.eup_r[q] = r[iu + eoff];
                    B.eup_g[q] = gc_raw(gc[iu + eoff]);
                }
#endif
                if (UR) {
                    const long ic = atc(g, cl.xl, y);
                    vload<float, V>(B.rv[q], r + ic);
                    if (xupd) {
                        vload<float, V>(B.xv[q], x + ic);
                        vload<float, V>(B.pv[q], pprev + ic);
                    }
                }
            }
        };
        // rows yb .. yb+RB-1 from a batch's loads (and the rows carried below)
        auto compute_batch = [&](const Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long ic = atc(g, cl.xl, y);
                float u2[V], gu[V];
                gc_unpack<V, FMT>(gu, B.gup[q]);
#pragma unroll
                for (int j = 0; j < V; ++j) u2[j] = fuse ? B.rup[q][j] + beta * B.up[q][j] : B.up[q][j];
                if (y + 1 >= g.ny) {  // reflection at the top edge: the row itself
#pragma unroll
                    for (int j = 0; j < V; ++j) {
                        u2[j] = cur[j];
                        gu[j] = gcur[j];
                    }
                }
                float lv, rvn, glv, grv;
#if TL_EDGE_PF
                auto col_val = [&](int) -> float { return ecur_p; };
                auto col_g = [&](int) -> float { return ecur_g; };
#else
                auto col_val = [&](int xx) -> float { return pnew(ic + (xx - cl.xl)); };
                auto col_g = [&](int xx) -> float { return gcol(ic + (xx - cl.xl)); };
#endif
                vneigh<float, V, TL_DPP && !UR>(g, cl, cur, col_val, lv, rvn);
                vneigh<float, V, TL_DPP && !UR>(g, cl, gcur, col_g, glv, grv);
                float wv[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const float gl = j > 0 ? gcur[j - 1] : glv;
                    float gr = j < V - 1 ? gcur[j + 1] : grv;
                    if (cl.xl + j + 1 >= g.nx) gr = gcur[j];
                    const float kxl = hx * (gl + gcur[j]), kxr = hx * (gcur[j] + gr);
                    const float kyd = hy * (gdn[j] + gcur[j]), kyu = hy * (gcur[j] + gu[j]);
                    wv[j] = vsmvp<float, V>(g, cl.xl, j, cur, lv, rvn, dn[j], u2[j], kxl, kxr, kyd, kyu);
                }
                if (y < cl.y1 && cl.active) {
                    const int nv = g.nx - cl.x0;
                    if (UR) {
                        float xn[V], rn[V];
#pragma unroll
                        for (int j = 0; j < V; ++j) {
                            // the two steps' updates in order, each rounded as if stored
                            xn[j] = xupd ? (B.xv[q][j] + alpha_prev * B.pv[q][j]) + alpha * cur[j] : 0.0f;
                            rn[j] = B.rv[q][j] - alpha * wv[j];
                            if (j < nv) acc[0] += (double)rn[j] * (double)rn[j];
                        }
                        if (xupd) vstore_n<float, V>(x + ic, xn, nv);
                        vstore_n<float, V>(r + ic, rn, nv);
                    } else {
                        if (fuse) vstore_n<float, V>(pout + ic, cur, nv);
#pragma unroll
                        for (int j = 0; j < V; ++j)
                            if (j < nv) acc[0] += (double)wv[j] * (double)cur[j];
                    }
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    dn[j] = cur[j];
                    cur[j] = u2[j];
                    gdn[j] = gcur[j];
                    gcur[j] = gu[j];
                }
#if TL_EDGE_PF
                // the row above becomes the current row (at the top edge the
                // clamped load read the current row itself)
                ecur_p = fuse ? B.eup_r[q] + beta * B.eup_p[q] : B.eup_p[q];
                ecur_g = gc_from_raw<FMT>(B.eup_g[q]);
#endif
            }
        };
        row_batches<PIPE, RB, Batch>(cl.y0, cl.y1, load_batch, compute_batch);
    }
    write_partials<1, TBV>(acc, part_out, g.nbmax);
}

// ----------------------- FP32 inner CG, one kernel per iteration (fused)
//
// Single-reduction CG: the two global reductions of an iteration (p.Ap before
// alpha, r.r before beta) are replaced by one, so that an iteration is one
// sweep over the grid.  Kernel t receives from kernel t-1 (or from k_cgf_dots
// after a start or an FP64 refresh) four sums for r = r(t-1) and p' = p(t-1):
//   gamma = (r, r), delta = (r, A r), eps = (r, A p'), pi = (p', A p')
// and derives, identically in every block,
//   beta = gamma / gamma(t-2)   (0 at t = 1)
//   (p, A p) = delta + 2 beta eps + beta^2 pi    for p = r + beta p'
//   alpha = gamma / (p, A p).
// (p, A p) is the expansion of (r + beta p', A (r + beta p')) and holds
// without assuming conjugacy, so it stays exact when an FP64 refresh replaces
// r and keeps p' (Chronopoulos and Gear's form, delta - beta gamma /
// alpha(t-1), assumes (r, A p') = -gamma / alpha(t-1), which the replacement
// breaks).  The kernel then applies the iteration: p = r + beta p',
// q = A p, r(t) = r - alpha q, x += alpha p (two iterations at a time, at even
// t, as in k_cgg), and accumulates the four sums of iteration t+1: (r(t),
// r(t)), (r(t), A r(t)), (r(t), q), (p, q).  A r(t) needs r(t) at the
// neighbours, hence q at the neighbours and p at distance 2: a thread walks
// its band keeping rows L-3 .. L in registers (p at L-2..L, r(t) at L-3..L-1,
// gc at L-3..L), takes the columns next to its 4 from the adjacent lanes by
// DPP wave shifts, and lanes 0 / 63 load the two columns outside the wave and
// compute q and r(t) at the first one themselves.
// Bytes per cell: reads r, p (4 each), gc (2), writes r(t), p (4 each): 18;
// at even t also reads and writes x: 26 (22 on average), against 34 for the
// two-kernel iteration (k_cgg PW + UR).  Needs nx % 4 == 0 (whole 4-column
// vectors per lane); other grids use k_cgg.

// one loaded row of the fused kernel: own 4 columns of r(t-1), p(t-1), x and
// the raw 16-bit coefficients; lanes 0 / 63 also the two columns outside
// the wave (lane 0: x0-2, x0-1; lane 63: x0+4, x0+5)
template <int RB>
struct FBatch {
    float r[RB][4], p[RB][4], x[RB][4];
    unsigned g[RB][2];
    float er[RB][2], ep[RB][2];
    unsigned eg[RB];
};

__device__ __forceinline__ unsigned dpp_up1u(unsigned v) { return (unsigned)dpp_up1((int)v); }
__device__ __forceinline__ unsigned dpp_down1u(unsigned v) { return (unsigned)dpp_down1((int)v); }
__device__ __forceinline__ float dpp_up1f(float v) { return __int_as_float(dpp_up1(__float_as_int(v))); }
__device__ __forceinline__ float dpp_down1f(float v) { return __int_as_float(dpp_down1(__float_as_int(v))); }

// 16-bit coefficient k (0 or 1) of a raw 32-bit word
template <int FMT>
__device__ __forceinline__ float gc_word(unsigned w, int k)
{
    return gc_from_raw<FMT>(k ? (w >> 16) : (w & 0xffffu));
}

// The FP32 operator of the fused kernel, flux form with a chain of fused
// multiply-adds: c + (kxr (c - r) + (kxl (c - l) + (kyu (c - u) + kyd (c - d))))
__device__ __forceinline__ float smvp5f(float c, float l, float r, float d, float u, float kxl, float kxr,
                                        float kyd, float kyu)
{
    float t = kyd * (c - d);
    t = __builtin_fmaf(kyu, c - u, t);
    t = __builtin_fmaf(kxl, c - l, t);
    t = __builtin_fmaf(kxr, c - r, t);
    return c + t;
}

// Row staging of the fused kernel through the LDS (DEPTH > 0).  Register
// loads keep about one row per wave in flight: the window of rows the kernel
// walks with takes most of its registers, so only 3 waves fit per SIMD and
// HBM latency, not bandwidth, bounds the sweep.  global_load_lds_dwordx4
// (gfx950) writes a row straight into the LDS without holding registers
// while in flight, so a wave can have DEPTH-1 rows requested ahead of the one
// it computes.  Per wave a ring of DEPTH slots of one row each; the DMA puts
// lane l's 16 bytes at slot + 16 l (4 bytes at + 4 l for the 4-byte form).
// The coefficient row (2 bytes per column) is fetched by lanes 0-31 with 16
// bytes each.  The columns outside the wave come by 4-byte gathers of lanes
// 0-4 (left: r, r, p, p at x0-2, x0-1, the gc word there) and 8-12 (right, at
// x0+256, x0+257).
struct alignas(16) FSlot {
    float r[256], p[256], x[256];   // own 4 columns of each lane, 16 bytes per lane
    unsigned g[128];                // coefficients, 8 bytes per lane
    unsigned e[16];                 // outside columns (words 0-4 left, 8-12 right)
};
constexpr unsigned FSLOT_P = 1024, FSLOT_X = 2048, FSLOT_G = 3072, FSLOT_E = 3584;
static_assert(sizeof(FSlot) == 3648 && offsetof(FSlot, p) == FSLOT_P && offsetof(FSlot, x) == FSLOT_X &&
                  offsetof(FSlot, g) == FSLOT_G && offsetof(FSlot, e) == FSLOT_E,
              "slot layout (the reads below use these offsets)");
#define TL_LDS(ptr) ((__attribute__((address_space(3))) void *)(ptr))

typedef float fv4_t __attribute__((ext_vector_type(4)));
typedef unsigned uv2_t __attribute__((ext_vector_type(2)));

// Read a staged row once at most N of this wave's vector-memory instructions
// are outstanding.  Vector memory counts complete in issue order on gfx9, and
// the N instructions issued after this row's DMAs are the later rows' DMAs
// (plus any stores, which only makes the wait longer).  The compiler does not
// see these LDS reads, so it inserts no full wait for the outstanding DMAs of
// the other slots; the block waits for its own reads before it ends.
template <int N, bool X>
__device__ __forceinline__ void fslot_read(unsigned a16, unsigned a8, unsigned ae, fv4_t &r, fv4_t &p, fv4_t &xv,
                                           uv2_t &gw, fv4_t &e4, unsigned &e1)
{
    if constexpr (X) {
        asm volatile(
            "s_waitcnt vmcnt(%[n])\n\t"
            "ds_read_b128 %[r], %[a16]\n\t"
            "ds_read_b128 %[p], %[a16] offset:1024\n\t"
            "ds_read_b128 %[x], %[a16] offset:2048\n\t"
            "ds_read_b64 %[g], %[a8]\n\t"
            "ds_read_b128 %[e], %[ae]\n\t"
            "ds_read_b32 %[e1], %[ae] offset:16\n\t"
            "s_waitcnt lgkmcnt(0)"
            : [r] "=&v"(r), [p] "=&v"(p), [x] "=&v"(xv), [g] "=&v"(gw), [e] "=&v"(e4), [e1] "=&v"(e1)
            : [a16] "v"(a16), [a8] "v"(a8), [ae] "v"(ae), [n] "i"(N)
            : "memory");
    } else {
        asm volatile(
            "s_waitcnt vmcnt(%[n])\n\t"
            "ds_read_b128 %[r], %[a16]\n\t"
            "ds_read_b128 %[p], %[a16] offset:1024\n\t"
            "ds_read_b64 %[g], %[a8]\n\t"
            "ds_read_b128 %[e], %[ae]\n\t"
            "ds_read_b32 %[e1], %[ae] offset:16\n\t"
            "s_waitcnt lgkmcnt(0)"
            : [r] "=&v"(r), [p] "=&v"(p), [g] "=&v"(gw), [e] "=&v"(e4), [e1] "=&v"(e1)
            : [a16] "v"(a16), [a8] "v"(a8), [ae] "v"(ae), [n] "i"(N)
            : "memory");
        xv = fv4_t{0.f, 0.f, 0.f, 0.f};
    }
}

// the wait for a row with rem later rows requested, G = 4 DMA instructions
// per row (5 when x is staged too)
template <int DEPTH>
__device__ __forceinline__ void fslot_wait_read(int rem, bool X, unsigned a16, unsigned a8, unsigned ae, fv4_t &r,
                                                fv4_t &p, fv4_t &xv, uv2_t &gw, fv4_t &e4, unsigned &e1)
{
    if (X) {
        if (DEPTH >= 4 && rem >= 3)
            fslot_read<15, true>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else if (DEPTH >= 3 && rem == 2)
            fslot_read<10, true>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else if (rem == 1)
            fslot_read<5, true>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else
            fslot_read<0, true>(a16, a8, ae, r, p, xv, gw, e4, e1);
    } else {
        if (DEPTH >= 4 && rem >= 3)
            fslot_read<12, false>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else if (DEPTH >= 3 && rem == 2)
            fslot_read<8, false>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else if (rem == 1)
            fslot_read<4, false>(a16, a8, ae, r, p, xv, gw, e4, e1);
        else
            fslot_read<0, false>(a16, a8, ae, r, p, xv, gw, e4, e1);
    }
}

template <int RB, int FMT, int DEPTH>
__global__ __launch_bounds__(TBV) TL_CGF_ATTR void k_cgf(Geo g, int ty, int t, const float *__restrict__ rin,
                                                         float *__restrict__ rout, const float *__restrict__ pin,
                                                         float *__restrict__ pout, float *__restrict__ x,
                                                         const tl_gc_t *__restrict__ gc, float hx, float hy,
                                                         Ctrl *__restrict__ ctrl, const double *__restrict__ sin,
                                                         unsigned nbin, double *__restrict__ sout)
{
    t = iter_of(t, ctrl);
    if (t > ctrl->stop_iter) return;
    const bool xupd = (t & 1) == 0;
    const VCol<4> cl = vcolumn<4>(g, ty);
    const int x0 = cl.x0, lane = cl.lane;
    // lanes at the wave's edges compute the column next to the wave (hcol:
    // x0-1 for lane 0, x0+4 for lane 63) from the two outside columns they
    // load (the other one: fcol); the grid's own edges clamp (zero flux)
    const bool eleft = lane == 0 && x0 > 0 && cl.active;
    const bool eright = lane == 63 && x0 + 4 < g.nx && cl.active;
    const bool eneed = eleft || eright;
    const long eoff = lane == 0 ? -2 : 4;     // first of the two outside columns
    const int hk = lane == 0 ? 1 : 0;         // which of the two is hcol
    const bool at_left = x0 == 0, at_right = x0 + 4 >= g.nx;

    auto load_batch = [&](FBatch<RB> &B, int yb) {
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int yy = min(max(yb + q, 0), g.ny - 1);
            const long i = at(g, cl.xl, yy);
            vload<float, 4>(B.r[q], rin + i);
            vload<float, 4>(B.p[q], pin + i);
            {
                // a cached load: the non-temporal form made the kernel 6-7% slower
                typedef unsigned uv2 __attribute__((ext_vector_type(2)));
                const uv2 w = *reinterpret_cast<const uv2 *>(gc + i);
                B.g[q][0] = w[0];
                B.g[q][1] = w[1];
            }
            if (xupd) vload<float, 4>(B.x[q], x + i);
            B.er[q][0] = B.er[q][1] = B.ep[q][0] = B.ep[q][1] = 0.f;
            B.eg[q] = 0u;
            if (eneed) {
                // the two outside columns start at an even element (x0 - 2 or
                // x0 + 4): one 8-byte load each of r and p, one 32-bit word of gc
                typedef float fv2 __attribute__((ext_vector_type(2)));
                const long ie = i + eoff;
                const fv2 er = *reinterpret_cast<const fv2 *>(rin + ie);
                const fv2 ep = *reinterpret_cast<const fv2 *>(pin + ie);
                B.er[q][0] = er[0];
                B.er[q][1] = er[1];
                B.ep[q][0] = ep[0];
                B.ep[q][1] = ep[1];
#if TL_GC16
                B.eg[q] = *reinterpret_cast<const unsigned *>(gc + ie);
#else
                B.eg[q] = (unsigned)gc_raw(gc[ie]) | ((unsigned)gc_raw(gc[ie + 1]) << 16);
#endif
            }
        }
    };

    // LDS staging (DEPTH > 0): this wave's ring, the per-lane sources of the
    // coefficient and outside-column gathers
    static_assert(DEPTH == 0 || RB == 1, "LDS staging works row by row");
    static_assert(DEPTH <= 4, "waits are spelled out for up to 3 rows ahead");
    __shared__ FSlot ring[DEPTH > 0 ? TBV / 64 : 1][DEPTH > 0 ? DEPTH : 1];
    const int xw = x0 - 4 * lane;                    // first column of the wave
    const int gcol = min(xw + 8 * lane, ((g.nx - 1) / 8) * 8);
    const bool glane = lane < 32;
    const bool elane = lane < 5 || (lane >= 8 && lane < 13);
    const char *ebase;
    long epitch;
    {
        const int k = lane & 7;
        int ec = lane < 8 ? (xw > 0 ? xw - 2 : 0) : (xw + 256 < g.nx ? xw + 256 : 0);
        if (k < 4) ec += k & 1;
        ebase = k < 2 ? (const char *)(rin + at(g, ec, 0))
                      : k < 4 ? (const char *)(pin + at(g, ec, 0)) : (const char *)(gc + at(g, ec, 0));
        epitch = g.pitch * (k < 4 ? (long)sizeof(float) : (long)sizeof(tl_gc_t));
    }
    FSlot *const myring = &ring[0][0] + (DEPTH > 0 ? (threadIdx.x >> 6) * DEPTH : 0);
    auto dma_row = [&](int s, int yb, bool withx) {
        const int yy = min(max(yb, 0), g.ny - 1);
        const long i = at(g, cl.xl, yy);
        FSlot &S = myring[s];
        __builtin_amdgcn_global_load_lds(rin + i, TL_LDS(S.r), 16, 0, 0);
        __builtin_amdgcn_global_load_lds(pin + i, TL_LDS(S.p), 16, 0, 0);
        if (withx) __builtin_amdgcn_global_load_lds(x + i, TL_LDS(S.x), 16, 0, 0);
        if (glane) __builtin_amdgcn_global_load_lds(gc + at(g, gcol, yy), TL_LDS(S.g), 16, 0, 0);
        if (elane) __builtin_amdgcn_global_load_lds(ebase + (long)yy * epitch, TL_LDS(S.e), 4, 0, 0);
    };

    // the band's first rows are requested before the block sums the previous
    // iteration's partials, so the reduction overlaps their latency
    FBatch<RB> B0, B1;
    if constexpr (DEPTH == 0) {
        load_batch(B0, cl.y0 - 2);
    } else {
#pragma unroll
        for (int k = 0; k < DEPTH - 1; ++k)
            if (cl.y0 - 2 + k <= cl.y1 + 1) dma_row(k, cl.y0 - 2 + k, xupd);
    }
    double s4[4];
    sum_partials<4, TBV>(sin, nbin, g.nbmax, s4);
    const double gam = s4[0];
    const int why = decide_why(ctrl, t - 1, gam);
    const double bd = t >= 2 ? gam / ctrl->rrs[(t - 2) & 1] : 0.0;
    const double pap = s4[1] + bd * (2.0 * s4[2] + bd * s4[3]);
    const double ad = gam / pap;
    const bool ok = pap > 0.0 && isfinite(ad);
    if (block_id() == 0 && threadIdx.x == 0) {
        ctrl->rrs[(t - 1) & 1] = gam;
        ctrl->rro = gam;
        ctrl->err = gam;
        ctrl->beta = bd;
        ctrl->iters = t - 1;
        record_stop(ctrl, t - 1, why);
        if (why == WHY_NONE) {
            ctrl->pw = pap;
            ctrl->alpha = ad;
            ctrl->alphas[t & 1] = ad;
            if (!ok) {  // non-positive curvature: stop before this iteration's update
                ctrl->status = TL_E_BREAKDOWN;
                ctrl->why = WHY_BREAKDOWN;
                ctrl->stop_iter = t - 1;
            }
        }
    }
    if (why != WHY_NONE || !ok) {  // every block derives the same decision
        // rows already requested into the LDS land before the wave ends
        if constexpr (DEPTH > 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        return;
    }
    const float beta = (float)bd, alpha = (float)ad;
    const float alpha_prev = xupd ? (float)ctrl->alphas[(t - 1) & 1] : 0.0f;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};     // gamma, delta, eps, pi of iteration t+1

    {
        const int y0 = cl.y0, y1 = cl.y1;  // y0 < ny: the grid has ceil(ny / ty) bands
        // The window, rows L-3 .. L (own 4 columns; *h: hcol, *f: fcol):
        //   p(t) at L-2, L-1, L; r(t-1) at L-1, L; r(t) at L-3, L-2, L-1;
        //   gc at L; face coefficients kx of rows L-2, L-1 (5 faces: left of
        //   column 0 .. right of column 3) and ky between rows L-3|L-2,
        //   L-2|L-1, L-1|L.  Faces are formed once per row and serve both
        //   A p(t) (row L-1) and, a row later, A r(t).
        float P0[4], P1[4], P2[4], P0h = 0.f, P1h = 0.f, P1f = 0.f, P2h = 0.f, P2f = 0.f;
        float R1[4], R2[4], R1h = 0.f, R2h = 0.f;
        float N0[4], N1[4], N2[4], N1h = 0.f, N2h = 0.f;
        float Gc[4], Gch = 0.f, Gcf = 0.f;           // gc of row L (own, hcol, fcol)
        unsigned Gw[2] = {0u, 0u};                   // its raw words (for the neighbour lanes)
        float KX1[5], KX2[5], KX1o = 0.f;            // kx of rows L-1, L-2; KX1o: fcol|hcol face of row L-1
        float KY0[4], KY1[4], KY2[4], KY1h = 0.f, KY2h = 0.f;  // ky L-3|L-2, L-2|L-1, L-1|L (hcol: 1, 2)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            P0[j] = P1[j] = P2[j] = R1[j] = R2[j] = N0[j] = N1[j] = N2[j] = 0.f;
            Gc[j] = KY0[j] = KY1[j] = KY2[j] = 0.f;
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) KX1[j] = KX2[j] = 0.f;

        // the value left / right of the own columns from the adjacent lane;
        // the edge lanes take their hcol value, the grid edges the cell itself
        auto nbr_l = [&](float src3, float c0, float hval) -> float {
            float v = dpp_up1f(src3);
            if (eleft) v = hval;
            if (at_left) v = c0;
            return v;
        };
        auto nbr_r = [&](float src0, float c3, float hval) -> float {
            float v = dpp_down1f(src0);
            if (eright) v = hval;
            if (at_right) v = c3;
            return v;
        };

        auto compute_batch = [&](const FBatch<RB> &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int L = yb + q;
                if (L > y1 + 1) break;
                // shift the window by one row
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    P0[j] = P1[j];
                    P1[j] = P2[j];
                    R1[j] = R2[j];
                    N0[j] = N1[j];
                    N1[j] = N2[j];
                    KY0[j] = KY1[j];
                    KY1[j] = KY2[j];
                }
#pragma unroll
                for (int j = 0; j < 5; ++j) KX2[j] = KX1[j];
                P0h = P1h;
                P1h = P2h;
                P1f = P2f;
                R1h = R2h;
                N1h = N2h;
                KY1h = KY2h;
                // faces of row L-1 from its coefficients (still in Gc) and row L's
                float gL[4];
                unsigned wL[2] = {B.g[q][0], B.g[q][1]};
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    gL[2 * k] = gc_word<FMT>(wL[k], 0);
                    gL[2 * k + 1] = gc_word<FMT>(wL[k], 1);
                }
                const float gLh = gc_word<FMT>(B.eg[q], hk), gLf = gc_word<FMT>(B.eg[q], 1 - hk);
                {
                    // kx of row L-1: neighbours' coefficients by DPP of the raw words
                    float gl = gc_word<FMT>(dpp_up1u(Gw[1]), 1), gr = gc_word<FMT>(dpp_down1u(Gw[0]), 0);
                    if (eleft) gl = Gch;
                    if (eright) gr = Gch;
                    KX1[0] = hx * (gl + Gc[0]);
#pragma unroll
                    for (int j = 1; j < 4; ++j) KX1[j] = hx * (Gc[j - 1] + Gc[j]);
                    KX1[4] = hx * (Gc[3] + gr);
                    KX1o = lane == 0 ? hx * (Gcf + Gch) : hx * (Gch + Gcf);
#pragma unroll
                    for (int j = 0; j < 4; ++j) KY2[j] = hy * (Gc[j] + gL[j]);
                    KY2h = hy * (Gch + gLh);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) Gc[j] = gL[j];
                Gch = gLh;
                Gcf = gLf;
                Gw[0] = wL[0];
                Gw[1] = wL[1];
                // row L: r(t-1) and p(t) = r + beta p(t-1)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    R2[j] = B.r[q][j];
                    P2[j] = __builtin_fmaf(beta, B.p[q][j], B.r[q][j]);
                }
                {
                    const float rh = B.er[q][hk], rf = B.er[q][1 - hk];
                    P2h = __builtin_fmaf(beta, B.ep[q][hk], rh);
                    P2f = __builtin_fmaf(beta, B.ep[q][1 - hk], rf);
                    R2h = rh;
                }
                if (L >= y0 && L < y1 && cl.active) {
                    const long i = at(g, x0, L);
                    vstore<float, 4>(pout + i, P2);
                    if (xupd) {
                        float xn[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            xn[j] = __builtin_fmaf(alpha, P2[j], __builtin_fmaf(alpha_prev, B.p[q][j], B.x[q][j]));
                        vstore<float, 4>(x + i, xn);
                    }
                }
                // q = A p(t) and r(t) = r(t-1) - alpha q on row L-1
                if (L >= y0) {
                    const int y = L - 1;
                    const float pl = nbr_l(P1[3], P1[0], P1h), pr = nbr_r(P1[0], P1[3], P1h);
                    const bool ybot = y <= 0, ytop = y >= g.ny - 1;
                    float qv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float c = P1[j];
                        const float l = j > 0 ? P1[j - 1] : pl;
                        const float r = j < 3 ? P1[j + 1] : pr;
                        const float d = ybot ? c : P0[j];
                        const float u = ytop ? c : P2[j];
                        qv[j] = smvp5f(c, l, r, d, u, KX1[j], KX1[j + 1], KY1[j], KY2[j]);
                        N2[j] = __builtin_fmaf(-alpha, qv[j], R1[j]);
                    }
                    {
                        // hcol of the edge lanes (fcol lies inside the grid when needed)
                        const float c = P1h;
                        const float own = lane == 0 ? P1[0] : P1[3];
                        const float l = lane == 0 ? P1f : own, r = lane == 0 ? own : P1f;
                        const float kxl = lane == 0 ? KX1o : KX1[4], kxr = lane == 0 ? KX1[0] : KX1o;
                        const float d = ybot ? c : P0h, u = ytop ? c : P2h;
                        const float qh = smvp5f(c, l, r, d, u, kxl, kxr, KY1h, KY2h);
                        N2h = __builtin_fmaf(-alpha, qh, R1h);
                    }
                    if (y >= y0 && y < y1 && cl.active) {
                        vstore<float, 4>(rout + at(g, x0, y), N2);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const double qd = (double)qv[j];
                            acc[2] += (double)N2[j] * qd;
                            acc[3] += (double)P1[j] * qd;
                        }
                    }
                }
                // A r(t) on row L-2, and the sums of r(t) there
                if (L >= y0 + 2) {
                    const int y = L - 2;
                    const float nl = nbr_l(N1[3], N1[0], N1h), nr = nbr_r(N1[0], N1[3], N1h);
                    const bool ybot = y <= 0, ytop = y >= g.ny - 1;
                    if (y < y1 && cl.active) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const float c = N1[j];
                            const float l = j > 0 ? N1[j - 1] : nl;
                            const float r = j < 3 ? N1[j + 1] : nr;
                            const float d = ybot ? c : N0[j];
                            const float u = ytop ? c : N2[j];
                            const float ar = smvp5f(c, l, r, d, u, KX2[j], KX2[j + 1], KY0[j], KY1[j]);
                            const double cd = (double)c;
                            acc[0] += cd * cd;
                            acc[1] += cd * (double)ar;
                        }
                    }
                }
            }
        };
        if constexpr (DEPTH == 0) {
            // rows y0-2 .. y1+1 in batches of RB, the next batch's loads issued
            // before the current batch is computed (the first one above)
            for (int yb = y0 - 2; yb < y1 + 2; yb += 2 * RB) {
                const bool more = yb + RB < y1 + 2;
                if (more) load_batch(B1, yb + RB);
                compute_batch(B0, yb);
                if (!more) break;
                if (yb + 2 * RB < y1 + 2) load_batch(B0, yb + 2 * RB);
                compute_batch(B1, yb + RB);
            }
        } else {
            // row L: request row L+DEPTH-1 into the slot row L-1 left, wait
            // for row L (the rows after it still in flight), compute it
            const unsigned ring0 = (unsigned)(size_t)TL_LDS(myring);
            const unsigned o16 = 16u * lane, o8 = FSLOT_G + 8u * lane, oe = FSLOT_E + (lane == 63 ? 32u : 0u);
            int sr = 0;  // slot of row L
            for (int L = y0 - 2; L <= y1 + 1; ++L) {
                const int ahead = L + DEPTH - 1;
                if (ahead <= y1 + 1) dma_row(sr == 0 ? DEPTH - 1 : sr - 1, ahead, xupd);
                const int rem = min(DEPTH - 1, y1 + 1 - L);  // rows requested after row L
                const unsigned base = ring0 + (unsigned)sr * (unsigned)sizeof(FSlot);
                fv4_t rv, pv, xv, e4;
                uv2_t gw;
                unsigned e1;
                fslot_wait_read<DEPTH>(rem, xupd, base + o16, base + o8, base + oe, rv, pv, xv, gw, e4, e1);
                FBatch<1> &B = B0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    B.r[0][j] = rv[j];
                    B.p[0][j] = pv[j];
                    B.x[0][j] = xv[j];
                }
                B.g[0][0] = gw[0];
                B.g[0][1] = gw[1];
                B.er[0][0] = eneed ? e4[0] : 0.f;
                B.er[0][1] = eneed ? e4[1] : 0.f;
                B.ep[0][0] = eneed ? e4[2] : 0.f;
                B.ep[0][1] = eneed ? e4[3] : 0.f;
                B.eg[0] = eneed ? e1 : 0u;
                compute_batch(B, L);
                sr = sr == DEPTH - 1 ? 0 : sr + 1;
            }
        }
    }
    write_partials<4, TBV>(acc, sout, g.nbmax);
}

// Two CG iterations per pair of sweeps (s-step CG with s = 2, monomial
// basis).  From r = r(t-1) and p' = p(t-1), with p = r + beta(t) p', the two
// iterations t and t+1 lie in the span of p, A p, A^2 p, r and A r:
//   alpha0 = gamma / (p, A p)                        gamma = (r, r)
//   r1 = r - alpha0 A p,        gamma1 = gamma - 2 alpha0 (r, A p) + alpha0^2 (A p, A p)
//   beta0 = gamma1 / gamma,     p1 = r1 + beta0 p,   A p1 = A r - alpha0 A^2 p + beta0 A p
//   (p1, A p1) = (r, A r) - 2 alpha0 (A p, A r) + alpha0^2 (A p, A^2 p)
//              + 2 beta0 ((r, A p) - alpha0 (A p, A p)) + beta0^2 (p, A p)
//   alpha1 = gamma1 / (p1, A p1)
//   x += alpha0 p + alpha1 p1,  r(t+1) = r1 - alpha1 A p1,  p(t+1) = p1.
// The first sweep (UPD false) forms p, A p, A^2 p and A r and sums the six
// products (p, Ap), (r, Ap), (Ap, Ap), (r, Ar), (Ap, Ar), (Ap, A^2 p); it
// stores nothing.  The second (UPD true) forms p and A p bit for bit as the
// first did, derives the scalars from the six sums, forms p1 on three rows
// and applies the stencil to it (TL_CGP_AP1; else A p1 from the basis as
// above), and writes r(t+1), p(t+1) and x, with (r(t+1), r(t+1)) for the
// next pair.  Bytes per cell for the two
// iterations: 10 (r, p', 16-bit gc) + 26 (also x; writes r, p, x) = 36, or
// 18 per iteration against 22 for k_cgf.  The walk is k_cgf's: rows L-3 .. L
// in registers (p at L-2..L, r at L-2..L, A p at L-3..L-1, A r at L-2, L-1),
// A p and A r at row L-1, A^2 p and the row's results at L-2; the edge lanes
// compute A p at the column next to the wave from the two outside columns.
// The first sweep of pair t applies the stop and refresh tests to
// gamma(t-1) as k_cgf does; when the pair's second step cannot be
// taken (gamma1 or (p1, A p1) not positive, or t + 1 beyond the iteration
// limit) the second sweep applies the first step only; when gamma1 already
// meets the stop or refresh test, the second step is a null step and the
// pair stops at t + 1 with p(t) kept.  Both sweeps run 3 waves per SIMD (768
// blocks, one round): the second needs 172 VGPRs unconstrained (2 waves);
// held to 168 it spills 4 and runs as fast as before its row windows were
// indexed by row (the first sweep: 147 VGPRs, 4% faster).
template <int FMT, bool UPD>
__global__ __launch_bounds__(TBV) __attribute__((amdgpu_waves_per_eu(3))) void k_cgp(Geo g, int ty, int t, const float *__restrict__ rin,
                                                         float *__restrict__ rout, const float *__restrict__ pin,
                                                         float *__restrict__ pout, float *__restrict__ x,
                                                         const tl_gc_t *__restrict__ gc, float hx, float hy,
                                                         Ctrl *__restrict__ ctrl, const double *__restrict__ sin,
                                                         unsigned nbin, double *__restrict__ sout)
{
    constexpr int NS = UPD ? 1 : 6;
    t = iter_of(t, ctrl);
    if (t > ctrl->stop_iter) return;
    const VCol<4> cl = vcolumn<4>(g, ty);
    const int x0 = cl.x0, lane = cl.lane;
    const bool eleft = lane == 0 && x0 > 0 && cl.active;
    const bool eright = lane == 63 && x0 + 4 < g.nx && cl.active;
    const bool eneed = eleft || eright;
    const long eoff = lane == 0 ? -2 : 4;
    const int hk = lane == 0 ? 1 : 0;
    const bool at_left = x0 == 0, at_right = x0 + 4 >= g.nx;

    auto load_batch = [&](FBatch<1> &B, int yb) {
        const int yy = min(max(yb, 0), g.ny - 1);
        const long i = at(g, cl.xl, yy);
        vload<float, 4>(B.r[0], rin + i);
        vload<float, 4>(B.p[0], pin + i);
        {
            typedef unsigned uv2 __attribute__((ext_vector_type(2)));
            const uv2 w = *reinterpret_cast<const uv2 *>(gc + i);
            B.g[0][0] = w[0];
            B.g[0][1] = w[1];
        }
        if constexpr (UPD) {  // x of the row the sweep finishes, two rows behind
            const int yx = min(max(yb - 2, 0), g.ny - 1);
            vload<float, 4>(B.x[0], x + at(g, cl.xl, yx));
        }
        B.er[0][0] = B.er[0][1] = B.ep[0][0] = B.ep[0][1] = 0.f;
        B.eg[0] = 0u;
        if (eneed) {
            typedef float fv2 __attribute__((ext_vector_type(2)));
            const long ie = i + eoff;
            const fv2 er = *reinterpret_cast<const fv2 *>(rin + ie);
            const fv2 ep = *reinterpret_cast<const fv2 *>(pin + ie);
            B.er[0][0] = er[0];
            B.er[0][1] = er[1];
            B.ep[0][0] = ep[0];
            B.ep[0][1] = ep[1];
            B.eg[0] = *reinterpret_cast<const unsigned *>(gc + ie);
        }
    };

    // the band's first row is requested before the block sums the previous
    // sweep's partials
    FBatch<1> Bs[2];
    load_batch(Bs[0], cl.y0 - 2);
    // scalars of the pair, identical in every block; keep: the second step is
    // a null step that leaves p(t) as the direction
    double bd, a0 = 0.0, b0 = 0.0, a1 = 0.0;
    bool keep = false;
    if constexpr (!UPD) {
        double gam;
        if (sin) {
            double s1[1];
            sum_partials<1, TBV>(sin, nbin, g.nbmax, s1);
            gam = s1[0];
        } else {
            gam = ctrl->rrs[(t - 1) & 1];  // after a start or an FP64 refresh
        }
        const int why = decide_why(ctrl, t - 1, gam);
        bd = t >= 2 ? gam / ctrl->rrs[(t - 2) & 1] : 0.0;
        if (block_id() == 0 && threadIdx.x == 0) {
            ctrl->rrs[(t - 1) & 1] = gam;
            ctrl->rro = gam;
            ctrl->err = gam;
            ctrl->beta = bd;
            ctrl->iters = t - 1;
            record_stop(ctrl, t - 1, why);
        }
        if (why != WHY_NONE) return;
    } else {
        double s6[6];
        sum_partials<6, TBV>(sin, nbin, g.nbmax, s6);
        const double gam = ctrl->rrs[(t - 1) & 1];
        bd = ctrl->beta;
        const double pap = s6[0], rap = s6[1], apap = s6[2], rar = s6[3], apar = s6[4], apa2p = s6[5];
        a0 = gam / pap;
        const bool ok = pap > 0.0 && isfinite(a0);
        const double g1 = gam - 2.0 * a0 * rap + a0 * a0 * apap;
        b0 = g1 / gam;
        const double p1ap1 =
            (rar - 2.0 * a0 * apar + a0 * a0 * apa2p) + 2.0 * b0 * (rap - a0 * apap) + b0 * b0 * pap;
        a1 = g1 / p1ap1;
        const bool cut = t + 1 > ctrl->maxiters;
        // the stop and refresh tests on gamma1 = (r(t), r(t)) as well: CG's
        // residual norm is not monotone, so a refresh the first step reaches
        // could be gone after the second.  Then the second step is a null
        // step (alpha1 = 0) that keeps p(t) and gamma(t-1), so that after the
        // refresh beta = gamma_new / gamma(t-1) as after a refresh at t, and
        // the pair stops at t + 1.
        const int why1 = ok && !cut && isfinite(g1) ? decide_why(ctrl, t, g1) : WHY_NONE;
        const bool early = why1 == WHY_REFRESH || why1 == WHY_CONVERGED;
        const bool two = ok && !cut && !early && g1 > 0.0 && p1ap1 > 0.0 && isfinite(a1);
        if (!two) {
            a1 = 0.0;
            b0 = 0.0;
        }
        keep = early;
        if (block_id() == 0 && threadIdx.x == 0) {
            ctrl->pw = pap;
            ctrl->alpha = two ? a1 : a0;
            ctrl->alphas[t & 1] = a0;
            ctrl->alphas[(t + 1) & 1] = a1;
            ctrl->rrs[t & 1] = early ? gam : g1 > 0.0 ? g1 : gam;
            ctrl->iters = cut ? t : t + 1;
            if (!ok) {  // non-positive curvature: stop before this pair's update
                ctrl->status = TL_E_BREAKDOWN;
                ctrl->why = WHY_BREAKDOWN;
                ctrl->iters = t - 1;
                ctrl->stop_iter = t - 1;
            } else if (cut) {
                record_stop(ctrl, t, WHY_MAXITERS);
            } else if (early) {
                record_stop(ctrl, t + 1, why1);
            }
        }
        if (!ok) return;
    }
    const float beta = (float)bd;
    const float fa0 = (float)a0, fb0 = (float)b0, fa1 = (float)a1;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;

    {
        const int y0 = cl.y0, y1 = cl.y1;
        // Row windows indexed by row: row m keeps slot (m - y0 + 2) % 3 of the
        // three-row windows and (m - y0 + 2) % 2 of the two-row ones while the
        // walk moves on, and the walk is unrolled six rows at a time, so every
        // index is a constant and no value moves between registers.  K = the
        // step's row L minus (y0 - 2), modulo 6.
        float P[3][4], Ph[3], Pf[2];     // p at L-2..L (own, hcol); fcol at L-1, L
        float R[3][4], Rh[2];            // r at L-2..L; hcol at L-1, L
        float N[3][4], Nh[2];            // A p at L-3..L-1; hcol at L-2, L-1
        float AR[2][4];                  // A r at L-2, L-1
        float Q[3][4], Qh[2];            // second sweep: p1 at L-3..L-1; hcol at L-2, L-1
        float Gc[2][4], Gch[2], Gcf[2];  // coefficients of rows L-1, L (own, hcol, fcol)
        unsigned Gw[2][2];               // their raw words (for the neighbour lanes)
        float KX[2][5];                  // kx of rows L-2, L-1 (5 faces)
        float KY[3][4], KYh[2];          // ky faces L-3|L-2, L-2|L-1, L-1|L; hcol of the last two
#pragma unroll
        for (int w = 0; w < 3; ++w) {
            Ph[w] = 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) P[w][j] = R[w][j] = N[w][j] = Q[w][j] = KY[w][j] = 0.f;
        }
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            Pf[w] = Rh[w] = Nh[w] = Qh[w] = Gch[w] = Gcf[w] = KYh[w] = 0.f;
            Gw[w][0] = Gw[w][1] = 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) AR[w][j] = Gc[w][j] = 0.f;
#pragma unroll
            for (int j = 0; j < 5; ++j) KX[w][j] = 0.f;
        }

        auto nbr_l = [&](float src3, float c0, float hval) -> float {
            float v = dpp_up1f(src3);
            if (eleft) v = hval;
            if (at_left) v = c0;
            return v;
        };
        auto nbr_r = [&](float src0, float c3, float hval) -> float {
            float v = dpp_down1f(src0);
            if (eright) v = hval;
            if (at_right) v = c3;
            return v;
        };

        auto compute_row = [&](const FBatch<1> &B, int L, auto kc) {
            constexpr int K = decltype(kc)::value;
            // slots: three-row windows (row L: a, L-1: b, L-2: c), two-row
            // windows (row L: e, L-1: f)
            constexpr int a = K % 3, b = (K + 2) % 3, c = (K + 1) % 3, e = K % 2, f = (K + 1) % 2;
            float gL[4];
            const unsigned wL[2] = {B.g[0][0], B.g[0][1]};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                gL[2 * k] = gc_word<FMT>(wL[k], 0);
                gL[2 * k + 1] = gc_word<FMT>(wL[k], 1);
            }
            const float gLh = gc_word<FMT>(B.eg[0], hk), gLf = gc_word<FMT>(B.eg[0], 1 - hk);
            float KXo;  // fcol|hcol face of row L-1
            {
                // faces of row L-1 (slot f) from its coefficients and row L's
                float gl = gc_word<FMT>(dpp_up1u(Gw[f][1]), 1), gr = gc_word<FMT>(dpp_down1u(Gw[f][0]), 0);
                if (eleft) gl = Gch[f];
                if (eright) gr = Gch[f];
                KX[f][0] = hx * (gl + Gc[f][0]);
#pragma unroll
                for (int j = 1; j < 4; ++j) KX[f][j] = hx * (Gc[f][j - 1] + Gc[f][j]);
                KX[f][4] = hx * (Gc[f][3] + gr);
                KXo = lane == 0 ? hx * (Gcf[f] + Gch[f]) : hx * (Gch[f] + Gcf[f]);
#pragma unroll
                for (int j = 0; j < 4; ++j) KY[a][j] = hy * (Gc[f][j] + gL[j]);
                KYh[e] = hy * (Gch[f] + gLh);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) Gc[e][j] = gL[j];
            Gch[e] = gLh;
            Gcf[e] = gLf;
            Gw[e][0] = wL[0];
            Gw[e][1] = wL[1];
            // row L: r and p = r + beta p'
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                R[a][j] = B.r[0][j];
                P[a][j] = __builtin_fmaf(beta, B.p[0][j], B.r[0][j]);
            }
            {
                const float rh = B.er[0][hk], rf = B.er[0][1 - hk];
                Ph[a] = __builtin_fmaf(beta, B.ep[0][hk], rh);
                Pf[e] = __builtin_fmaf(beta, B.ep[0][1 - hk], rf);
                Rh[e] = rh;
            }
            // A p and A r on row L-1.  The rows outside the grid were loaded
            // clamped, so p and r there equal the edge row's, bit for bit:
            // the zero-flux neighbours need no select here (A^2 p below, from
            // A p of rows outside the grid, does)
            if (L >= y0) {
                const float pl = nbr_l(P[b][3], P[b][0], Ph[b]), pr = nbr_r(P[b][0], P[b][3], Ph[b]);
                const float rl = nbr_l(R[b][3], R[b][0], Rh[f]), rr = nbr_r(R[b][0], R[b][3], Rh[f]);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float l = j > 0 ? P[b][j - 1] : pl;
                    const float r = j < 3 ? P[b][j + 1] : pr;
                    N[b][j] = smvp5f(P[b][j], l, r, P[c][j], P[a][j], KX[f][j], KX[f][j + 1], KY[b][j], KY[a][j]);
                    if (!UPD || !TL_CGP_AP1) {
                        const float lr = j > 0 ? R[b][j - 1] : rl;
                        const float rrr = j < 3 ? R[b][j + 1] : rr;
                        AR[f][j] =
                            smvp5f(R[b][j], lr, rrr, R[c][j], R[a][j], KX[f][j], KX[f][j + 1], KY[b][j], KY[a][j]);
                    }
                }
                {
                    // A p at the column next to the wave (edge lanes)
                    const float own = lane == 0 ? P[b][0] : P[b][3];
                    const float l = lane == 0 ? Pf[f] : own, r = lane == 0 ? own : Pf[f];
                    const float kxl = lane == 0 ? KXo : KX[f][4], kxr = lane == 0 ? KX[f][0] : KXo;
                    Nh[f] = smvp5f(Ph[b], l, r, Ph[c], Ph[a], kxl, kxr, KYh[f], KYh[e]);
                }
                if constexpr (UPD && TL_CGP_AP1) {
                    // p1 = r1 + beta0 p on row L-1 (and the next column), r1 = r - alpha0 A p
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        Q[b][j] = __builtin_fmaf(fb0, P[b][j], __builtin_fmaf(-fa0, N[b][j], R[b][j]));
                    Qh[f] = __builtin_fmaf(fb0, Ph[b], __builtin_fmaf(-fa0, Nh[f], Rh[f]));
                }
            }
            // A^2 p on row L-2 and the row's sums or updates (row L-3: slot a
            // of the three-row windows that end at L-1)
            if (L >= y0 + 2) {
                const int y = L - 2;
                if (y < y1 && cl.active) {
                    const bool ybot = y <= 0, ytop = y >= g.ny - 1;
                    const float nl = nbr_l(N[c][3], N[c][0], Nh[e]), nr = nbr_r(N[c][0], N[c][3], Nh[e]);
                    float ql = 0.f, qr = 0.f;
                    if constexpr (UPD && TL_CGP_AP1) {
                        ql = nbr_l(Q[c][3], Q[c][0], Qh[e]);
                        qr = nbr_r(Q[c][0], Q[c][3], Qh[e]);
                    }
                    (void)ql;
                    (void)qr;
                    float o_r[4], o_p[4], o_x[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if constexpr (UPD && TL_CGP_AP1) {
                            // the second step as the fused kernel takes a step:
                            // A p1 by the stencil of p1
                            const float cq = Q[c][j];
                            const float l = j > 0 ? Q[c][j - 1] : ql;
                            const float r = j < 3 ? Q[c][j + 1] : qr;
                            const float d = ybot ? cq : Q[a][j];
                            const float u = ytop ? cq : Q[b][j];
                            const float ap1 = smvp5f(cq, l, r, d, u, KX[e][j], KX[e][j + 1], KY[c][j], KY[b][j]);
                            const float r1 = __builtin_fmaf(-fa0, N[c][j], R[c][j]);
                            o_x[j] = __builtin_fmaf(fa1, cq, __builtin_fmaf(fa0, P[c][j], B.x[0][j]));
                            o_r[j] = __builtin_fmaf(-fa1, ap1, r1);
                            o_p[j] = keep ? P[c][j] : cq;
                            acc[0] = __builtin_fma((double)o_r[j], (double)o_r[j], acc[0]);
                            continue;
                        }
                        const float cn = N[c][j];
                        const float l = j > 0 ? N[c][j - 1] : nl;
                        const float r = j < 3 ? N[c][j + 1] : nr;
                        const float d = ybot ? cn : N[a][j];
                        const float u = ytop ? cn : N[b][j];
                        const float a2p = smvp5f(cn, l, r, d, u, KX[e][j], KX[e][j + 1], KY[c][j], KY[b][j]);
                        const float pv = P[c][j], apv = N[c][j], rv = R[c][j], arv = AR[e][j];
                        if constexpr (!UPD) {
                            // a product of two FP32 values is exact in FP64, so
                            // the fused multiply-add rounds as acc + a * b does
                            const double ap = (double)apv, rd = (double)rv, ard = (double)arv;
                            acc[0] = __builtin_fma((double)pv, ap, acc[0]);
                            acc[1] = __builtin_fma(rd, ap, acc[1]);
                            acc[2] = __builtin_fma(ap, ap, acc[2]);
                            acc[3] = __builtin_fma(rd, ard, acc[3]);
                            acc[4] = __builtin_fma(ap, ard, acc[4]);
                            acc[5] = __builtin_fma(ap, (double)a2p, acc[5]);
                            (void)o_r;
                            (void)o_p;
                            (void)o_x;
                        } else {
                            const float r1 = __builtin_fmaf(-fa0, apv, rv);
                            const float p1 = __builtin_fmaf(fb0, pv, r1);
                            const float ar1 = __builtin_fmaf(-fa0, a2p, arv);
                            const float ap1 = __builtin_fmaf(fb0, apv, ar1);
                            o_x[j] = __builtin_fmaf(fa1, p1, __builtin_fmaf(fa0, pv, B.x[0][j]));
                            o_r[j] = __builtin_fmaf(-fa1, ap1, r1);
                            o_p[j] = keep ? pv : p1;
                            acc[0] = __builtin_fma((double)o_r[j], (double)o_r[j], acc[0]);
                        }
                    }
                    if constexpr (UPD) {
                        const long i = at(g, x0, y);
                        vstore<float, 4>(rout + i, o_r);
                        vstore<float, 4>(pout + i, o_p);
                        vstore<float, 4>(x + i, o_x);
                    }
                }
            }
        };
        // rows y0-2 .. y1+1, the next row's loads issued before the current
        // row is computed (the first one above; two rows ahead in the first
        // sweep measured no faster: it is bound by its arithmetic)
        auto step = [&](auto kc, int L) -> bool {
            constexpr int K = decltype(kc)::value;
            if (L > y1 + 1) return false;
            if (L + 1 <= y1 + 1) load_batch(Bs[(K + 1) % 2], L + 1);
            compute_row(Bs[K % 2], L, kc);
            return true;
        };
        for (int L = y0 - 2; L <= y1 + 1; L += 6) {
            if (!step(std::integral_constant<int, 0>{}, L)) break;
            if (!step(std::integral_constant<int, 1>{}, L + 1)) break;
            if (!step(std::integral_constant<int, 2>{}, L + 2)) break;
            if (!step(std::integral_constant<int, 3>{}, L + 3)) break;
            if (!step(std::integral_constant<int, 4>{}, L + 4)) break;
            if (!step(std::integral_constant<int, 5>{}, L + 5)) break;
        }
    }
    write_partials<NS, TBV>(acc, sout, g.nbmax);
}

// The four sums k_cgf needs at its first iteration and after an FP64 refresh,
// for the residual r and the kept direction p: (r, r), (r, A r), (r, A p),
// (p, A p), with the same FP32 operator (flux form, clamped neighbours).  One
// thread per cell of a column band; runs once per start or refresh.
template <int FMT>
__global__ __launch_bounds__(TB) void k_cgf_dots(Geo g, const float *__restrict__ r, const float *__restrict__ p,
                                                 const tl_gc_t *__restrict__ gc, float hx, float hy,
                                                 double *__restrict__ part)
{
    const Col c = column(g);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (c.active) {
        const int x = c.x;
        const int xl = max(x - 1, 0), xr = min(x + 1, g.nx - 1);
        auto G = [&](int xx, int yy) -> float { return gc_from_raw<FMT>(gc_raw(gc[at(g, xx, yy)])); };
        for (int y = c.y0; y < c.y1; ++y) {
            const int yd = max(y - 1, 0), yu = min(y + 1, g.ny - 1);
            const float gm = G(x, y), gL = G(xl, y), gR = G(xr, y), gD = G(x, yd), gU = G(x, yu);
            const float kxl = hx * (gL + gm), kxr = hx * (gm + gR), kyd = hy * (gD + gm), kyu = hy * (gm + gU);
            const long i = at(g, x, y);
            const float rc = r[i], pc = p[i];
            const float ar = smvp5f(rc, r[at(g, xl, y)], r[at(g, xr, y)], r[at(g, x, yd)], r[at(g, x, yu)], kxl,
                                    kxr, kyd, kyu);
            const float ap = smvp5f(pc, p[at(g, xl, y)], p[at(g, xr, y)], p[at(g, x, yd)], p[at(g, x, yu)], kxl,
                                    kxr, kyd, kyu);
            const double rd = (double)rc;
            acc[0] += rd * rd;
            acc[1] += rd * (double)ar;
            acc[2] += rd * (double)ap;
            acc[3] += (double)pc * (double)ap;
        }
    }
    write_partials<4>(acc, part, g.nbmax);
}

// ---------------------------------------------- residuals and summaries

// residual!, kernels.jl:227-232: r = u0 - A u.  Optional outputs: r (FP64),
// r32 and p32 (FP32 copies for the mixed-precision CG), partial sums of r^2.
// Skipped when t > stop_iter (t = 0 always runs).
__global__ __launch_bounds__(TB) void k_residual(Geo g, int t, const double *__restrict__ u0,
                                                 const double *__restrict__ u, const double *__restrict__ kx,
                                                 const double *__restrict__ ky, double *__restrict__ r,
                                                 float *__restrict__ r32, float *__restrict__ p32,
                                                 const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (t > ctrl->stop_iter) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        Win<double> win;
        win.start(g, c, u, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<double> st;
            double b[R];
            load_stn(st, g, c, yb, u, kx, ky);
            load_rows(b, g, c, yb, u0);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double rv = b[q] - win.apply(st, q);
                if (yb + q < c.y1) {
                    const long i = at(g, c.x, yb + q);
                    if (r) r[i] = rv;
                    if (r32) {
                        const float rf = (float)rv;
                        r32[i] = rf;
                        if (p32) p32[i] = rf;
                    }
                    acc[0] += rv * rv;
                }
                win.advance(st, q);
            }
        }
    }
    if (part) write_partials<1>(acc, part, g.nbmax);
}

// partial sums of (b - A x)^2
__global__ __launch_bounds__(TB) void k_resid_norm(Geo g, const double *__restrict__ b,
                                                   const double *__restrict__ xs, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double *__restrict__ part)
{
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        Win<double> win;
        win.start(g, c, xs, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<double> st;
            double bb[R];
            load_stn(st, g, c, yb, xs, kx, ky);
            load_rows(bb, g, c, yb, b);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double rv = bb[q] - win.apply(st, q);
                if (yb + q < c.y1) acc[0] += rv * rv;
                win.advance(st, q);
            }
        }
    }
    write_partials<1>(acc, part, g.nbmax);
}

// dst = A src over the interior
__global__ __launch_bounds__(TB) void k_apply(Geo g, const double *__restrict__ a, const double *__restrict__ kx,
                                              const double *__restrict__ ky, double *__restrict__ o)
{
    const Col c = column(g);
    if (!c.active) return;
    Win<double> win;
    win.start(g, c, a, ky);
    for (int yb = c.y0; yb < c.y1; yb += R) {
        Stn<double> st;
        load_stn(st, g, c, yb, a, kx, ky);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const double v = win.apply(st, q);
            if (yb + q < c.y1) o[at(g, c.x, yb + q)] = v;
            win.advance(st, q);
        }
    }
}

// finalise!, kernels.jl:239-242: energy = u / density
__global__ __launch_bounds__(TB) void k_finalise(Geo g, const double *__restrict__ u, const double *__restrict__ rho,
                                                 double *__restrict__ e)
{
    const Col c = column(g);
    if (!c.active) return;
    for (int yb = c.y0; yb < c.y1; yb += R) {
        double uv[R], dv[R];
        load_rows(uv, g, c, yb, u);
        load_rows(dv, g, c, yb, rho);
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (yb + q < c.y1) e[at(g, c.x, yb + q)] = uv[q] / dv[q];
    }
}

// fieldsummary, kernels.jl:119-133: partials of vol, mass, ie, temp
__global__ __launch_bounds__(TB) void k_field_summary(Geo g, const double *__restrict__ rho,
                                                      const double *__restrict__ e, const double *__restrict__ u,
                                                      double vol, double *__restrict__ part)
{
    const Col c = column(g);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (c.active) {
        for (int yb = c.y0; yb < c.y1; yb += R) {
            double dv[R], ev[R], uv[R];
            load_rows(dv, g, c, yb, rho);
            load_rows(ev, g, c, yb, e);
            load_rows(uv, g, c, yb, u);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (yb + q < c.y1) {
                    const double m = vol * dv[q];
                    acc[0] += vol;
                    acc[1] += m;
                    acc[2] += m * ev[q];
                    acc[3] += m * uv[q];
                }
            }
        }
    }
    write_partials<4>(acc, part, g.nbmax);
}

// partial sums of v^2 (skipped when t > stop_iter; t = 0 always runs)
__global__ __launch_bounds__(TB) void k_sumsq(Geo g, int t, const double *__restrict__ v,
                                              const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (t > ctrl->stop_iter) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        for (int yb = c.y0; yb < c.y1; yb += R) {
            double vv[R];
            load_rows(vv, g, c, yb, v);
#pragma unroll
            for (int q = 0; q < R; ++q)
                if (yb + q < c.y1) acc[0] += vv[q] * vv[q];
        }
    }
    write_partials<1>(acc, part, g.nbmax);
}

// ---------------------------------------------------------------- halo

struct HaloArgs {
    double *f[8];
    int n;
};

// updateface!, kernels.jl:191-210 (reflective, index-corrected): for each
// field, depth layers on the four faces over the interior span.
__global__ __launch_bounds__(TB) void k_halo(Geo g, HaloArgs h, int depth)
{
    const int t = blockIdx.x * TB + threadIdx.x;
    const int fi = blockIdx.y;
    const int layer = blockIdx.z;  // 0 .. depth-1
    if (fi >= h.n || layer >= depth) return;
    double *a = h.f[fi];
    const long o = g.off;
    if (t < g.ny) {
        const long row = o + (long)t * g.pitch;
        a[row - 1 - layer] = a[row + layer];
        a[row + g.nx + layer] = a[row + g.nx - 1 - layer];
    }
    if (t < g.nx) {
        a[o + (long)(-1 - layer) * g.pitch + t] = a[o + (long)layer * g.pitch + t];
        a[o + (long)(g.ny + layer) * g.pitch + t] = a[o + (long)(g.ny - 1 - layer) * g.pitch + t];
    }
}

// ------------------------------------------------------------ Chebyshev

// Cheby.init!, Cheby.jl:64-81: w = A u; r = u0 - w; p = r/theta; u' = u + p;
// partials of bb = sum u0^2 (u' is a separate buffer: the stencil reads u's
// neighbours)
__global__ __launch_bounds__(TB) void k_cheby_init(Geo g, const double *__restrict__ uin, double *__restrict__ uout,
                                                   const double *__restrict__ u0, double *__restrict__ p,
                                                   double *__restrict__ r, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double theta,
                                                   double *__restrict__ part)
{
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        Win<double> win;
        win.start(g, c, uin, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<double> st;
            double b[R];
            load_stn(st, g, c, yb, uin, kx, ky);
            load_rows(b, g, c, yb, u0);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double rv = b[q] - win.apply(st, q);
                const double pv = rv / theta;
                if (yb + q < c.y1) {
                    const long i = at(g, c.x, yb + q);
                    r[i] = rv;
                    p[i] = pv;
                    uout[i] = win.cur + pv;
                    acc[0] += b[q] * b[q];
                }
                win.advance(st, q);
            }
        }
    }
    write_partials<1>(acc, part, g.nbmax);
}

// ----------------------------------------------------------------- PPCG

// CG.ur! without its norm (PPCG.jl:67, value unused) fused with
// PPCG.init!(chunk, hd) (PPCG.jl:97-100): u += alpha p; r -= alpha w; sd = r/theta
__global__ __launch_bounds__(TB) void k_ppcg_ur_sd(Geo g, int t, double *__restrict__ u,
                                                   const double *__restrict__ p, double *__restrict__ r,
                                                   const double *__restrict__ w, double *__restrict__ sd,
                                                   double theta, Ctrl *__restrict__ ctrl,
                                                   const double *__restrict__ pw, unsigned nbw,
                                                   const double *__restrict__ prprev, unsigned nbp)
{
    if (t > ctrl->stop_iter) return;
    double alpha;
    if (!cg_alpha(g, t, ctrl, pw, nbw, prprev, nbp, nullptr, &alpha)) return;
    const Col c = column(g);
    if (!c.active) return;
    for (int yb = c.y0; yb < c.y1; yb += R) {
        double uv[R], pv[R], rv[R], wv[R];
        load_rows(uv, g, c, yb, u);
        load_rows(pv, g, c, yb, p);
        load_rows(rv, g, c, yb, r);
        load_rows(wv, g, c, yb, w);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (yb + q < c.y1) {
                const long i = at(g, c.x, yb + q);
                u[i] = uv[q] + alpha * pv[q];
                const double rn = rv[q] - alpha * wv[q];
                r[i] = rn;
                sd[i] = rn / theta;
            }
        }
    }
}

// ------------------------------- vectorised FP64 Chebyshev / PPCG / Jacobi
//
// The three stencil-update kernels of the other FP64 solvers on 16-byte
// vectors (2 doubles per lane), in the row-batch scheme of k_cg_pwv: a lane
// walks its column pair up a band keeping the rows below / at the current
// row in registers, takes the columns left / right of its pair from the
// adjacent lanes (lanes 0 and 63 load them with the row loads), and the next
// batch's loads are issued before the current one is computed.  Per cell the
// arithmetic is the reference's, operand for operand (smvp5's association;
// Jacobi's as k_jacobi below), so results do not depend on the vector width.
//   VS_CHEBY  (Cheby.mainstep!, Cheby.jl:84-106):
//             r = u0 - A u; p = a p + b r; u' = u + p; [sum r^2 when calc2norm]
//   VS_PPCG   (PPCG inner step, PPCG.jl:75-84 in the two-phase form: sd is
//             ping-ponged): r -= A sd; u += sd; sd' = a sd + b r; [sum r^2 at
//             the last inner step, PPCG.jl:88, consumed by the next CG.p!]
//   VS_JACOBI (Jacobi.iterate!, Jacobi.jl:59-80):
//             u' = (u0 + sum K u_nbr) / (1 + sum K); sum |u' - u|
// `a_in` is the stencil input (u or sd), `a_out` its successor (u' or sd');
// f1 / f2 are u0 / p (Chebyshev), r / u (PPCG), u0 / unused (Jacobi).
enum VsOp : int { VS_CHEBY = 0, VS_PPCG = 1, VS_JACOBI = 2 };

template <int OP, int RB, bool PIPE>
__global__ __launch_bounds__(TBV) void k_vs64(Geo g, int ty, int t, const double *__restrict__ a_in,
                                              double *__restrict__ a_out, const double *__restrict__ kx,
                                              const double *__restrict__ ky, double *__restrict__ f1,
                                              double *__restrict__ f2, double ca, double cb, int sums,
                                              const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    constexpr int V = 2;
    typedef double T;
    if (t > ctrl->stop_iter) return;
    const VCol<V> cl = vcolumn<V>(g, ty);
    double acc[1] = {0.0};
    if (cl.y0 < g.ny) {
        T dn[V], cur[V], kyd[V];
        {
            const long i0 = at(g, cl.xl, cl.y0), id = at(g, cl.xl, max(cl.y0 - 1, 0));
            vload<T, V>(cur, a_in + i0);
            vload<T, V>(dn, a_in + id);
            vload<T, V>(kyd, ky + i0);
        }
        const bool eneed = (cl.lane == 0 && cl.xl > 0) || (cl.lane == 63 && cl.xl + V < g.nx);
        const long eoff = cl.lane == 0 ? -1 : V;
        T ecur = (T)0;
        if (eneed) ecur = a_in[at(g, cl.xl, cl.y0) + eoff];
        struct Batch {
            T up[RB][V], kxv[RB][V], kyu[RB][V], e1[RB][V], e2[RB][V];
            T eup[RB], ekx[RB];
        };
        auto load_batch = [&](Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long ic = atc(g, cl.xl, y), iu = atc(g, cl.xl, y + 1);
                vload<T, V>(B.up[q], a_in + iu);
                vload<T, V>(B.kxv[q], kx + ic);
                vload<T, V>(B.kyu[q], ky + at(g, cl.xl, min(y + 1, g.ny)));
                vload<T, V>(B.e1[q], f1 + ic);
                if (OP != VS_JACOBI) vload<T, V>(B.e2[q], f2 + ic);
                B.eup[q] = B.ekx[q] = (T)0;
                if (eneed) B.eup[q] = a_in[iu + eoff];
                if (cl.lane == 63) B.ekx[q] = kx[ic + V];
            }
        };
        auto compute_batch = [&](const Batch &B, int yb) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int y = yb + q;
                const long ic = atc(g, cl.xl, y);
                T u2[V];
#pragma unroll
                for (int j = 0; j < V; ++j) u2[j] = y + 1 >= g.ny ? cur[j] : B.up[q][j];
                T lv, rv;
                auto col_val = [&](int) -> T { return ecur; };
                vneigh<T, V>(g, cl, cur, col_val, lv, rv);
                T kxn = lane_down1<false>(B.kxv[q][0]);
                if (cl.lane == 63) kxn = B.ekx[q];
                T o1[V], o2[V], o3[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxl = B.kxv[q][j], kxr = j < V - 1 ? B.kxv[q][j + 1] : kxn;
                    const T kyu = B.kyu[q][j];
                    if constexpr (OP == VS_JACOBI) {
                        const T l = j > 0 ? cur[j - 1] : lv;
                        T r = j < V - 1 ? cur[j + 1] : rv;
                        if (cl.xl + j + 1 >= g.nx) r = cur[j];
                        double num = B.e1[q][j] + kxr * r;
                        num = num + kxl * l;
                        num = num + kyu * u2[j];
                        num = num + kyd[j] * dn[j];
                        double den = 1.0 + kxl;
                        den = den + kxr;
                        den = den + kyd[j];
                        den = den + kyu;
                        o1[j] = num / den;
                    } else {
                        const T w = vsmvp<T, V>(g, cl.xl, j, cur, lv, rv, dn[j], u2[j], kxl, kxr, kyd[j], kyu);
                        if constexpr (OP == VS_CHEBY) {
                            const T rr = B.e1[q][j] - w;        // u0 - A u
                            const T pv = ca * B.e2[q][j] + cb * rr;
                            o1[j] = cur[j] + pv;                // u'
                            o2[j] = pv;                         // p
                            o3[j] = rr;
                        } else {
                            const T rr = B.e1[q][j] - w;        // r - A sd
                            o1[j] = ca * cur[j] + cb * rr;      // sd'
                            o2[j] = rr;                         // r
                            o3[j] = B.e2[q][j] + cur[j];        // u + sd
                        }
                    }
                }
                if (y < cl.y1 && cl.active) {
                    const int nv = g.nx - cl.x0;
                    vstore_n<T, V>(a_out + ic, o1, nv);
                    if constexpr (OP == VS_CHEBY) {
                        vstore_n<T, V>(f2 + ic, o2, nv);
#pragma unroll
                        for (int j = 0; j < V; ++j)
                            if (j < nv) acc[0] += o3[j] * o3[j];
                    } else if constexpr (OP == VS_PPCG) {
                        vstore_n<T, V>(f1 + ic, o2, nv);
                        vstore_n<T, V>(f2 + ic, o3, nv);
#pragma unroll
                        for (int j = 0; j < V; ++j)
                            if (j < nv) acc[0] += o2[j] * o2[j];
                    } else {
#pragma unroll
                        for (int j = 0; j < V; ++j)
                            if (j < nv) acc[0] += fabs(o1[j] - cur[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    dn[j] = cur[j];
                    cur[j] = u2[j];
                    kyd[j] = B.kyu[q][j];
                }
                ecur = B.eup[q];
            }
        };
        row_batches<PIPE, RB, Batch>(cl.y0, cl.y1, load_batch, compute_batch);
    }
    if (sums) write_partials<1, TBV>(acc, part, g.nbmax);
}

// Two PPCG inner steps (PPCG.jl:78-83, sd ping-ponged) in one sweep.  The
// inner steps have no reduction between them, so step 2 can follow step 1
// a row behind in the same pass: a thread walks its band keeping sd0 at rows
// L-2..L and sd1 at rows L-3..L-1 (with the column next to the wave for the
// lanes at its edges), and at row L
//   step 1 at row M = L-1:  r1 = r0 - A sd0,  sd1 = a1 sd0 + b1 r1
//   step 2 at row N = L-2:  r2 = r1 - A sd1,  sd2 = a2 sd1 + b2 r2,
//                           u2 = (u0 + sd0) + sd1
// with the single step's operations and association (k_vs64, VS_PPCG), so
// the result is bit-identical to two launches of it.  Lanes 0 / 63 compute
// step 1 also at the column next to the wave (hcol), from sd0 one column
// further (fcol), as the neighbouring wave does for its own column.  Reads
// sd0, r0, u, kx, ky once and writes sd2, r2, u2 once: 64 B/cell for two
// steps instead of 128.  r is read at neighbours' rows and columns by other
// threads, so it is not updated in place (rin -> rout); u is (own cell only).
__global__ __launch_bounds__(TBV) void k_ppcg2(Geo g, int ty, int t, const double *__restrict__ sdin,
                                               double *__restrict__ sdout, const double *__restrict__ rin,
                                               double *__restrict__ rout, double *__restrict__ u,
                                               const double *__restrict__ kx, const double *__restrict__ ky,
                                               double a1, double b1, double a2, double b2, int sums,
                                               const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    constexpr int V = 2;
    typedef double T;
    if (t > ctrl->stop_iter) return;
    const VCol<V> cl = vcolumn<V>(g, ty);
    const int xl = cl.xl, lane = cl.lane;
    double acc[1] = {0.0};
    if (cl.y0 < g.ny) {
        const bool eleft = lane == 0 && xl > 0;
        const bool eright = lane == 63 && xl + V < g.nx;
        const bool edge = eleft || eright;
        const int hc = lane == 0 ? xl - 1 : xl + V;                       // hcol
        const int hcc = edge ? hc : xl;                                     // a valid column to load
        const int fcc = eleft ? xl - 2 : eright ? min(xl + V + 1, g.nx - 1) : xl;
        // one row's loads: sd0 at row L (own, hcol, fcol); r0, kx at row L-1;
        // ky at row L (the face below it); u at row L-2
        struct Row {
            T s[V], r[V], kx[V], ky[V], u[V];
            T es, ef, er, ekx0, ekx1, eky;
        };
        auto load_row = [&](Row &B, int L) {
            const int ys = min(max(L, 0), g.ny - 1), yr = min(max(L - 1, 0), g.ny - 1);
            const int yk = min(max(L, 0), g.ny), yu = min(max(L - 2, 0), g.ny - 1);
            vload<T, V, TL_NT_2STEP_LD != 0>(B.s, sdin + at(g, xl, ys));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.r, rin + at(g, xl, yr));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.kx, kx + at(g, xl, yr));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.ky, ky + at(g, xl, yk));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.u, u + at(g, xl, yu));
            B.es = B.ef = B.er = B.ekx1 = B.eky = (T)0;
            // lane 63's right face of its last column, as the single step loads it
            B.ekx0 = lane == 63 ? kx[at(g, xl, yr) + V] : (T)0;
            if (edge) {
                B.es = sdin[at(g, hcc, ys)];
                B.ef = sdin[at(g, fcc, ys)];
                B.er = rin[at(g, hcc, yr)];
                B.eky = ky[at(g, hcc, yk)];
                if (lane == 0) B.ekx0 = kx[at(g, xl - 1, yr)];               // left face of hcol
                else B.ekx1 = kx[at(g, min(xl + V + 1, g.nx), yr)];          // right face of hcol
            }
        };
        // the window at row L, before step 1 at row M: sd0 at rows N, M (own,
        // hcol) and fcol at M; kx at N, ky at N and M; sd1 at rows N-1, N
        // (own; hcol at N) and r1 at N, from the two previous rows
        T S0m[V], S0c[V], SHm = 0, SHc = 0, EFc = 0;
        T KXm[V], EKXm = 0, KYm[V], KYc[V], EKYc = 0;
        T S1b[V], S1c[V], H1c = 0, R1c[V];
#pragma unroll
        for (int j = 0; j < V; ++j) S0m[j] = S0c[j] = KXm[j] = KYm[j] = KYc[j] = S1b[j] = S1c[j] = R1c[j] = (T)0;
        const int y0 = cl.y0, y1 = cl.y1;
        auto step = [&](const Row &B, int L) {
            const int M = L - 1, N = L - 2;
            // step 1 at row M (own columns, and hcol on the edge lanes)
            T s1[V], r1[V], h1 = 0;
            {
                const bool ybot = M <= 0, ytop = M + 1 >= g.ny;
                T lv = lane_up1<false>(S0c[V - 1]), rv = lane_down1<false>(S0c[0]);
                if (lane == 0) lv = xl > 0 ? SHc : S0c[0];
                if (lane == 63) rv = xl + V < g.nx ? SHc : S0c[V - 1];
                if (xl == 0) lv = S0c[0];
                T kxn = lane_down1<false>(B.kx[0]);
                if (lane == 63) kxn = B.ekx0;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxl = B.kx[j], kxr = j < V - 1 ? B.kx[j + 1] : kxn;
                    const T dn = ybot ? S0c[j] : S0m[j], up = ytop ? S0c[j] : B.s[j];
                    const T w = vsmvp<T, V>(g, xl, j, S0c, lv, rv, dn, up, kxl, kxr, KYc[j], B.ky[j]);
                    r1[j] = B.r[j] - w;
                    s1[j] = a1 * S0c[j] + b1 * r1[j];
                }
                if (edge) {
                    const T c = SHc;
                    T l, r, kxl, kxr;
                    if (lane == 0) {
                        l = hc == 0 ? c : EFc;
                        r = S0c[0];
                        kxl = B.ekx0;
                        kxr = B.kx[0];
                    } else {
                        l = S0c[V - 1];
                        r = hc + 1 >= g.nx ? c : EFc;
                        kxl = B.ekx0;
                        kxr = B.ekx1;
                    }
                    const T dn = ybot ? c : SHm, up = ytop ? c : B.es;
                    const T w = smvp5(c, l, r, dn, up, kxl, kxr, EKYc, B.eky);
                    const T rh = B.er - w;
                    h1 = a1 * c + b1 * rh;
                }
            }
            // step 2 at row N (own columns), from sd1 at rows N-1, N, M
            if (L >= y0 + 2 && N < y1) {
                const bool ybot = N <= 0, ytop = N + 1 >= g.ny;
                T lv = lane_up1<false>(S1c[V - 1]), rv = lane_down1<false>(S1c[0]);
                if (lane == 0) lv = xl > 0 ? H1c : S1c[0];
                if (lane == 63) rv = xl + V < g.nx ? H1c : S1c[V - 1];
                if (xl == 0) lv = S1c[0];
                T kxn = lane_down1<false>(KXm[0]);
                if (lane == 63) kxn = EKXm;
                T o1[V], o2[V], o3[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxl = KXm[j], kxr = j < V - 1 ? KXm[j + 1] : kxn;
                    const T dn = ybot ? S1c[j] : S1b[j], up = ytop ? S1c[j] : s1[j];
                    const T w = vsmvp<T, V>(g, xl, j, S1c, lv, rv, dn, up, kxl, kxr, KYm[j], KYc[j]);
                    const T rr = R1c[j] - w;
                    o1[j] = a2 * S1c[j] + b2 * rr;          // sd2
                    o2[j] = rr;                             // r2
                    o3[j] = (B.u[j] + S0m[j]) + S1c[j];     // u2
                }
                if (cl.active) {
                    const long ic = at(g, xl, N);
                    const int nv = g.nx - cl.x0;
                    vstore_n<T, V, TL_NT_2STEP_ST != 0>(sdout + ic, o1, nv);
                    vstore_n<T, V, TL_NT_2STEP_ST != 0>(rout + ic, o2, nv);
                    vstore_n<T, V, TL_NT_2STEP_ST != 0>(u + ic, o3, nv);
#pragma unroll
                    for (int j = 0; j < V; ++j)
                        if (j < nv) acc[0] += o2[j] * o2[j];
                }
            }
            // shift the window by one row
#pragma unroll
            for (int j = 0; j < V; ++j) {
                S0m[j] = S0c[j];
                S0c[j] = B.s[j];
                KXm[j] = B.kx[j];
                KYm[j] = KYc[j];
                KYc[j] = B.ky[j];
                S1b[j] = S1c[j];
                S1c[j] = s1[j];
                R1c[j] = r1[j];
            }
            SHm = SHc;
            SHc = B.es;
            EFc = B.ef;
            EKXm = B.ekx0;
            EKYc = B.eky;
            H1c = h1;
        };
        // rows y0-2 .. y1+1, the next row's loads issued before the current
        // row is computed
        Row B0, B1;
        load_row(B0, y0 - 2);
        for (int L = y0 - 2; L <= y1 + 1; L += 2) {
            const bool more = L + 1 <= y1 + 1;
            if (more) load_row(B1, L + 1);
            step(B0, L);
            if (!more) break;
            if (L + 2 <= y1 + 1) load_row(B0, L + 2);
            step(B1, L + 1);
        }
    }
    if (sums) write_partials<1, TBV>(acc, part, g.nbmax);
}

// Two Chebyshev steps (Cheby.iterate!, Cheby.jl:93-107) in one sweep, as
// k_ppcg2 does for PPCG: step 1 at row M = L-1, step 2 at row N = L-2,
//   step 1:  r1 = u0 - A u,   p1 = a1 p0 + b1 r1,  u1 = u + p1
//   step 2:  r2 = u0 - A u1,  p2 = a2 p1 + b2 r2,  u2 = u1 + p2
// with the single step's operations and association (k_vs64, VS_CHEBY), so
// the result is bit-identical to two launches of it; with norm, the partial
// sums of r2^2 as the single step forms them for its residual norm.  u and p
// are read at the column next to the wave by its edge lanes, so both move to
// other buffers (u_in -> u_out, p_in -> p_out).  Reads u, u0, p, kx, ky once
// and writes u2, p2 once: 56 B/cell for two steps instead of 112.
__global__ __launch_bounds__(TBV) void k_cheby2(Geo g, int ty, int t, const double *__restrict__ uin,
                                                double *__restrict__ uout, const double *__restrict__ u0,
                                                const double *__restrict__ pin, double *__restrict__ pout,
                                                const double *__restrict__ kx, const double *__restrict__ ky,
                                                double a1, double b1, double a2, double b2, int norm,
                                                const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    constexpr int V = 2;
    typedef double T;
    if (t > ctrl->stop_iter) return;
    const VCol<V> cl = vcolumn<V>(g, ty);
    const int xl = cl.xl, lane = cl.lane;
    double acc[1] = {0.0};
    if (cl.y0 < g.ny) {
        const bool eleft = lane == 0 && xl > 0;
        const bool eright = lane == 63 && xl + V < g.nx;
        const bool edge = eleft || eright;
        const int hc = lane == 0 ? xl - 1 : xl + V;                       // hcol
        const int hcc = edge ? hc : xl;                                     // a valid column to load
        const int fcc = eleft ? xl - 2 : eright ? min(xl + V + 1, g.nx - 1) : xl;
        // one row's loads: u at row L (own, hcol, fcol); u0, p0, kx at row
        // L-1; ky at row L (the face below it)
        struct Row {
            T u[V], u0[V], p[V], kx[V], ky[V];
            T eu, ef, eu0, ep, ekx0, ekx1, eky;
        };
        auto load_row = [&](Row &B, int L) {
            const int ys = min(max(L, 0), g.ny - 1), yr = min(max(L - 1, 0), g.ny - 1);
            const int yk = min(max(L, 0), g.ny);
            vload<T, V, TL_NT_2STEP_LD != 0>(B.u, uin + at(g, xl, ys));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.u0, u0 + at(g, xl, yr));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.p, pin + at(g, xl, yr));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.kx, kx + at(g, xl, yr));
            vload<T, V, TL_NT_2STEP_LD != 0>(B.ky, ky + at(g, xl, yk));
            B.eu = B.ef = B.eu0 = B.ep = B.ekx1 = B.eky = (T)0;
            // lane 63's right face of its last column, as the single step loads it
            B.ekx0 = lane == 63 ? kx[at(g, xl, yr) + V] : (T)0;
            if (edge) {
                B.eu = uin[at(g, hcc, ys)];
                B.ef = uin[at(g, fcc, ys)];
                B.eu0 = u0[at(g, hcc, yr)];
                B.ep = pin[at(g, hcc, yr)];
                B.eky = ky[at(g, hcc, yk)];
                if (lane == 0) B.ekx0 = kx[at(g, xl - 1, yr)];               // left face of hcol
                else B.ekx1 = kx[at(g, min(xl + V + 1, g.nx), yr)];          // right face of hcol
            }
        };
        // the window at row L, before step 1 at row M: u at rows N, M (own,
        // hcol) and fcol at M; u0, kx at N, ky at N and M; u1 at rows N-1, N
        // (own; hcol at N) and p1 at N, from the two previous rows
        T Um[V], Uc[V], UHm = 0, UHc = 0, EFc = 0;
        T U0m[V], KXm[V], EKXm = 0, KYm[V], KYc[V], EKYc = 0;
        T V1b[V], V1c[V], H1c = 0, P1c[V];
#pragma unroll
        for (int j = 0; j < V; ++j) Um[j] = Uc[j] = U0m[j] = KXm[j] = KYm[j] = KYc[j] = V1b[j] = V1c[j] = P1c[j] = (T)0;
        const int y0 = cl.y0, y1 = cl.y1;
        auto step = [&](const Row &B, int L) {
            const int M = L - 1, N = L - 2;
            // step 1 at row M (own columns, and hcol on the edge lanes)
            T v1[V], p1[V], h1 = 0;
            {
                const bool ybot = M <= 0, ytop = M + 1 >= g.ny;
                T lv = lane_up1<false>(Uc[V - 1]), rv = lane_down1<false>(Uc[0]);
                if (lane == 0) lv = xl > 0 ? UHc : Uc[0];
                if (lane == 63) rv = xl + V < g.nx ? UHc : Uc[V - 1];
                if (xl == 0) lv = Uc[0];
                T kxn = lane_down1<false>(B.kx[0]);
                if (lane == 63) kxn = B.ekx0;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxl = B.kx[j], kxr = j < V - 1 ? B.kx[j + 1] : kxn;
                    const T dn = ybot ? Uc[j] : Um[j], up = ytop ? Uc[j] : B.u[j];
                    const T w = vsmvp<T, V>(g, xl, j, Uc, lv, rv, dn, up, kxl, kxr, KYc[j], B.ky[j]);
                    const T rr = B.u0[j] - w;
                    p1[j] = a1 * B.p[j] + b1 * rr;
                    v1[j] = Uc[j] + p1[j];
                }
                if (edge) {
                    const T c = UHc;
                    T l, r, kxl, kxr;
                    if (lane == 0) {
                        l = hc == 0 ? c : EFc;
                        r = Uc[0];
                        kxl = B.ekx0;
                        kxr = B.kx[0];
                    } else {
                        l = Uc[V - 1];
                        r = hc + 1 >= g.nx ? c : EFc;
                        kxl = B.ekx0;
                        kxr = B.ekx1;
                    }
                    const T dn = ybot ? c : UHm, up = ytop ? c : B.eu;
                    const T w = smvp5(c, l, r, dn, up, kxl, kxr, EKYc, B.eky);
                    const T rr = B.eu0 - w;
                    const T ph = a1 * B.ep + b1 * rr;
                    h1 = c + ph;
                }
            }
            // step 2 at row N (own columns), from u1 at rows N-1, N, M
            if (L >= y0 + 2 && N < y1) {
                const bool ybot = N <= 0, ytop = N + 1 >= g.ny;
                T lv = lane_up1<false>(V1c[V - 1]), rv = lane_down1<false>(V1c[0]);
                if (lane == 0) lv = xl > 0 ? H1c : V1c[0];
                if (lane == 63) rv = xl + V < g.nx ? H1c : V1c[V - 1];
                if (xl == 0) lv = V1c[0];
                T kxn = lane_down1<false>(KXm[0]);
                if (lane == 63) kxn = EKXm;
                T o1[V], o2[V], o3[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const T kxl = KXm[j], kxr = j < V - 1 ? KXm[j + 1] : kxn;
                    const T dn = ybot ? V1c[j] : V1b[j], up = ytop ? V1c[j] : v1[j];
                    const T w = vsmvp<T, V>(g, xl, j, V1c, lv, rv, dn, up, kxl, kxr, KYm[j], KYc[j]);
                    const T rr = U0m[j] - w;                // u0 - A u1
                    const T pv = a2 * P1c[j] + b2 * rr;
                    o1[j] = V1c[j] + pv;                    // u2
                    o2[j] = pv;                             // p2
                    o3[j] = rr;
                }
                if (cl.active) {
                    const long ic = at(g, xl, N);
                    const int nv = g.nx - cl.x0;
                    vstore_n<T, V, TL_NT_2STEP_ST != 0>(uout + ic, o1, nv);
                    vstore_n<T, V, TL_NT_2STEP_ST != 0>(pout + ic, o2, nv);
#pragma unroll
                    for (int j = 0; j < V; ++j)
                        if (j < nv) acc[0] += o3[j] * o3[j];
                }
            }
            // shift the window by one row
#pragma unroll
            for (int j = 0; j < V; ++j) {
                Um[j] = Uc[j];
                Uc[j] = B.u[j];
                U0m[j] = B.u0[j];
                KXm[j] = B.kx[j];
                KYm[j] = KYc[j];
                KYc[j] = B.ky[j];
                V1b[j] = V1c[j];
                V1c[j] = v1[j];
                P1c[j] = p1[j];
            }
            UHm = UHc;
            UHc = B.eu;
            EFc = B.ef;
            EKXm = B.ekx0;
            EKYc = B.eky;
            H1c = h1;
        };
        // rows y0-2 .. y1+1, the next row's loads issued before the current
        // row is computed
        Row B0, B1;
        load_row(B0, y0 - 2);
        for (int L = y0 - 2; L <= y1 + 1; L += 2) {
            const bool more = L + 1 <= y1 + 1;
            if (more) load_row(B1, L + 1);
            step(B0, L);
            if (!more) break;
            if (L + 2 <= y1 + 1) load_row(B0, L + 2);
            step(B1, L + 1);
        }
    }
    if (norm) write_partials<1, TBV>(acc, part, g.nbmax);
}

// --------------------------------------------------------------- Jacobi

// Jacobi.iterate!, Jacobi.jl:59-80 (u' in a separate buffer, the reference's
// r = copy of u): u' = (u0 + sum K u_nbr) / (1 + sum K); partials of
// err = sum |u' - u|
__global__ __launch_bounds__(TB) void k_jacobi(Geo g, int t, const double *__restrict__ uin,
                                               double *__restrict__ uout, const double *__restrict__ u0,
                                               const double *__restrict__ kx, const double *__restrict__ ky,
                                               const Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (t > ctrl->stop_iter) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        Win<double> win;
        win.start(g, c, uin, ky);
        for (int yb = c.y0; yb < c.y1; yb += R) {
            Stn<double> st;
            double b[R];
            load_stn(st, g, c, yb, uin, kx, ky);
            load_rows(b, g, c, yb, u0);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double num = b[q] + st.kxr[q] * st.rt[q];
                num = num + st.kxl[q] * st.lf[q];
                num = num + st.kyu[q] * st.up[q];
                num = num + win.kyd * win.dn;
                double den = 1.0 + st.kxl[q];
                den = den + st.kxr[q];
                den = den + win.kyd;
                den = den + st.kyu[q];
                const double un = num / den;
                if (yb + q < c.y1) {
                    uout[at(g, c.x, yb + q)] = un;
                    acc[0] += fabs(un - win.cur);
                }
                win.advance(st, q);
            }
        }
    }
    write_partials<1>(acc, part, g.nbmax);
}

// ------------------------------------------------------ mixed precision

// cell coefficient of the inner CG over the whole allocation:
// gc = 1/w with w = density (coefficient 1) or 1/density (CG.jl:56-58),
// rounded to FP32, multiplied by scale (a power of two: exact) and rounded to
// the storage format, to nearest with ties to even; cells without a density
// (padding) get 0
template <int FMT>
__device__ __forceinline__ tl_gc_t gc_store(float f)
{
#if TL_GC16
    if constexpr (FMT == GC_F16) {
        return __builtin_bit_cast(unsigned short, (_Float16)f);
    } else {
        unsigned u = __float_as_uint(f);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (tl_gc_t)(u >> 16);
    }
#else
    return f;
#endif
}

__device__ __forceinline__ float gc_value(double d, int coefficient)
{
    return coefficient == 2 ? (float)d : (d != 0.0 ? (float)(1.0 / d) : 0.0f);
}

template <int FMT>
__global__ __launch_bounds__(TB) void k_gcell(const double *__restrict__ rho, tl_gc_t *__restrict__ gc, size_t n,
                                              int coefficient, float scale)
{
    const size_t i = (size_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    gc[i] = gc_store<FMT>(gc_value(rho[i], coefficient) * scale);
}

// per-block minimum and maximum of gc over the cells with a density (grid-
// stride over the allocation; TL_RANGE_BLOCKS blocks): out[2b], out[2b + 1]
__global__ __launch_bounds__(TB) void k_gc_range(const double *__restrict__ rho, size_t n, int coefficient,
                                                 float *__restrict__ out)
{
    float lo = INFINITY, hi = 0.0f;
    for (size_t i = (size_t)blockIdx.x * TB + threadIdx.x; i < n; i += (size_t)gridDim.x * TB) {
        const double d = rho[i];
        if (d != 0.0) {
            const float f = gc_value(d, coefficient);
            lo = fminf(lo, f);
            hi = fmaxf(hi, f);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    __shared__ float slo[TB / 64], shi[TB / 64];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        slo[wave] = lo;
        shi[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TB / 64; ++w) {
            lo = fminf(lo, slo[w]);
            hi = fmaxf(hi, shi[w]);
        }
        out[2 * blockIdx.x] = lo;
        out[2 * blockIdx.x + 1] = hi;
    }
}

// FP64 solution accumulation of the mixed-precision CG after iteration T:
// u += x; x = 0.  x holds the updates up to the last even iteration; for odd
// T the update of iteration T, alpha(T) p(T), is added first, as the FP32 UR
// kernel would have.
__global__ __launch_bounds__(TB) void k_ir_accumulate(Geo g, double *__restrict__ u, float *__restrict__ x,
                                                      const float *__restrict__ pT, const Ctrl *__restrict__ ctrl,
                                                      int T)
{
    const Col c = column(g);
    if (!c.active) return;
    const bool pend = (T & 1) && pT;
    const float aT = pend ? (float)ctrl->alphas[T & 1] : 0.0f;
    for (int yb = c.y0; yb < c.y1; yb += R) {
        double uv[R];
        float xv[R], pv[R];
        load_rows(uv, g, c, yb, u);
        load_rows(xv, g, c, yb, x);
        if (pend) load_rows(pv, g, c, yb, pT);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (yb + q < c.y1) {
                const long i = at(g, c.x, yb + q);
                const float xf = pend ? xv[q] + aT * pv[q] : xv[q];
                u[i] = uv[q] + (double)xf;
                x[i] = 0.0f;
            }
        }
    }
}

// ------------------------------------------------------------- launchers

static inline dim3 sgrid(const Geo &g) { return dim3(g.gx, g.gy); }
static inline dim3 fgrid(const Geo &g) { return dim3((g.X + TB - 1) / TB, g.Y); }
static inline unsigned nbh(const tl_ctx *c) { return (unsigned)(c->g.gx * c->g.gy); }
static inline unsigned nb4(const tl_ctx *c) { return (unsigned)(c->g.gx4 * c->g.gy4); }
static inline unsigned nb4p(const tl_ctx *c) { return (unsigned)(c->g.gx4 * c->g.gyp4); }
static inline unsigned nb2(const tl_ctx *c) { return (unsigned)(c->g.gx2 * c->g.gy2); }
static inline size_t stride(const tl_ctx *c) { return (size_t)c->g.nbmax; }
// partial-sum arrays: [pw | pr0 | pr1 | pm0 pm1 pm2 pm3], nbmax doubles each
static inline double *part_w(tl_ctx *c) { return c->partials; }
static inline double *part_r(tl_ctx *c, int t) { return c->partials + (size_t)(1 + (t & 1)) * stride(c); }
static inline double *part_m(tl_ctx *c) { return c->partials + (size_t)3 * stride(c); }
static inline unsigned &nbr(tl_ctx *c, int t) { return c->nb_r[t & 1]; }
// p of iteration t >= 1 of the fused CG lives in buffer (t-1) & 1
static inline double *p64(tl_ctx *c, int t) { return ((t - 1) & 1) ? c->p_alt : c->f[TL_F_P]; }
static inline float *p32(tl_ctx *c, int t) { return ((t - 1) & 1) ? c->p32b : c->p32; }

// kernel argument for iteration t: t itself, or, in a launch captured into a
// hipGraph batch whose first iteration is tref, -(t - tref) (the kernel adds
// the replay's first iteration, Ctrl::tg)
static inline int targ(int t, int tref) { return tref > 0 ? -(t - tref) : t; }

// event pair around one launch when this kernel is being profiled (never
// inside a graph capture)
static bool prof_start(tl_ctx *c, int kind)
{
    if (c->capturing || !(c->prof_mask & (1u << kind))) return false;
    if (c->prof_seq[kind]++ % c->prof_period != 0) return false;
    if (c->prof_used + 2 > c->prof_ev.size()) {
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
        c->prof_ev.push_back(e0);
        c->prof_ev.push_back(e1);
        c->prof_tag.push_back(kind);
    }
    c->prof_tag[c->prof_used / 2] = kind;
    return hipEventRecord(c->prof_ev[c->prof_used], c->stream) == hipSuccess;
}

static void prof_stop(tl_ctx *c)
{
    (void)hipEventRecord(c->prof_ev[c->prof_used + 1], c->stream);
    c->prof_used += 2;
}

static hipError_t fin(tl_ctx *c, const double *part, unsigned nb, int n, int slot, int kind, int t)
{
    k_fin<<<1, TB, 0, c->stream>>>(c->ctrl, part, nb, (unsigned)stride(c), n, slot, kind, t);
    return hipGetLastError();
}

hipError_t launch_init_ctrl(tl_ctx *c, int mode, int maxiters, double eps, double rel_tol, double ir_delta,
                            int presteps, int errorswitch, double epslim)
{
    k_init_ctrl<<<1, 64, 0, c->stream>>>(c->ctrl, mode, maxiters, eps, rel_tol, ir_delta, presteps, errorswitch,
                                        epslim);
    return hipGetLastError();
}

hipError_t launch_set_mode(tl_ctx *c, int mode, int maxiters)
{
    k_set_mode<<<1, 64, 0, c->stream>>>(c->ctrl, mode, maxiters);
    return hipGetLastError();
}

hipError_t launch_set_state(tl_ctx *c, const tl_state *states_d, int nstates)
{
    k_set_state<<<fgrid(c->g), TB, 0, c->stream>>>(c->g, states_d, nstates, c->xmin, c->ymin, c->dx, c->dy,
                                                   c->f[TL_F_DENSITY], c->f[TL_F_ENERGY0], c->f[TL_F_U]);
    return hipGetLastError();
}

hipError_t launch_copy_full(tl_ctx *c, const double *a, double *b)
{
    k_copy_full<<<fgrid(c->g), TB, 0, c->stream>>>(c->g, a, b);
    return hipGetLastError();
}

hipError_t launch_copy_vec(tl_ctx *c, const double *a, double *b)
{
    const size_t n2 = c->g.elems / 2;
    k_copy_vec<<<(unsigned)((n2 + TB - 1) / TB), TB, 0, c->stream>>>((const double2 *)a, (double2 *)b, n2);
    return hipGetLastError();
}

hipError_t launch_copy_energy(tl_ctx *c) { return launch_copy_full(c, c->f[TL_F_ENERGY0], c->f[TL_F_ENERGY]); }

hipError_t launch_init_coef(tl_ctx *c, double rx, double ry, int coefficient, int what)
{
    k_init_coef<<<fgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_DENSITY], c->f[TL_F_ENERGY], c->f[TL_F_U],
                                                   c->f[TL_F_P], c->f[TL_F_R], c->f[TL_F_U0], c->f[TL_F_KX],
                                                   c->f[TL_F_KY], rx, ry, coefficient == 2 ? 1 : 0, what);
    return hipGetLastError();
}

hipError_t launch_cg_init_resid(tl_ctx *c, const double *guess)
{
    k_cg_init_resid<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], guess, c->f[TL_F_U0], c->f[TL_F_P],
                                                       c->f[TL_F_R], c->f[TL_F_W], c->f[TL_F_KX], c->f[TL_F_KY],
                                                       part_r(c, 0), part_m(c));
    nbr(c, 0) = nbh(c);
    c->nb_m = nbh(c);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 2, 0, FK_INIT, 0) : e;
}

hipError_t launch_resid_norm_slot(tl_ctx *c, const double *b, const double *x, int slot)
{
    k_resid_norm<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, b, x, c->f[TL_F_KX], c->f[TL_F_KY], part_m(c));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 1, slot, FK_SUMS, 0) : e;
}

hipError_t launch_cg_w64(tl_ctx *c, int t)
{
    const bool prof = prof_start(c, TL_K_CG_W64);
    k_cg_w<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, c->f[TL_F_P], c->f[TL_F_KX], c->f[TL_F_KY],
                                                       c->f[TL_F_W], c->ctrl, part_w(c));
    c->nb_w = nbh(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_cg_ur64(tl_ctx *c, int t)
{
    const bool prof = prof_start(c, TL_K_CG_UR64);
    k_cg_ur<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, c->f[TL_F_U], c->f[TL_F_P], c->f[TL_F_R],
                                                        c->f[TL_F_W], c->ctrl, part_w(c), c->nb_w, part_r(c, t - 1),
                                                        nbr(c, t - 1), part_r(c, t), c->hist_a);
    nbr(c, t) = nbh(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_cg_p64(tl_ctx *c, int t)
{
    const bool prof = prof_start(c, TL_K_CG_P64);
    k_cg_p<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, c->f[TL_F_P], c->f[TL_F_R], c->ctrl,
                                                       part_r(c, t - 1), nbr(c, t - 1), part_r(c, t), nbr(c, t),
                                                       c->hist_b);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

// fused CG, FP64: PW(t) then UR(t)
hipError_t launch_cg_pw64(tl_ctx *c, int t, int tref)
{
    const bool prof = prof_start(c, TL_K_CG_W64);
    k_cg_pwv<double, 2, TL_RV_PW64, TL_PIPE_PW64 != 0><<<nb2(c), TBV, 0, c->stream>>>(
        c->g, c->g.ty2, targ(t, tref), t >= 2 ? p64(c, t - 1) : nullptr, p64(c, t), c->f[TL_F_R], c->f[TL_F_KX], c->f[TL_F_KY],
        c->f[TL_F_W], c->ctrl, part_w(c), t >= 2 ? part_r(c, t - 1) : nullptr, t >= 2 ? nbr(c, t - 1) : 0,
        c->hist_b);
    c->nb_w = nb2(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_cg_urv64(tl_ctx *c, int t, int tref)
{
    const bool prof = prof_start(c, TL_K_CG_UR64);
    k_cg_urv<double, 2, TL_RV_UR64, TL_PIPE_UR64 != 0><<<nb2(c), TBV, 0, c->stream>>>(
        c->g, c->g.ty2, targ(t, tref), c->cg_u ? c->cg_u : c->f[TL_F_U], p64(c, t), c->f[TL_F_R], c->f[TL_F_W], c->ctrl,
        part_w(c), c->nb_w,
        part_r(c, t), c->hist_a);
    nbr(c, t) = nb2(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

// the CG.p! of the last iteration T of a fused FP64 CG solve (p_{T+1} in the
// buffer of p_T, as the unfused solver leaves it); returns that buffer
hipError_t launch_cg_pfinal64(tl_ctx *c, int T)
{
    k_cg_p<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, T, p64(c, T), c->f[TL_F_R], c->ctrl, part_r(c, T - 1),
                                                       nbr(c, T - 1), part_r(c, T), nbr(c, T), c->hist_b);
    return hipGetLastError();
}

hipError_t launch_cg_w32(tl_ctx *c, int t, int tref)
{
    const bool prof = prof_start(c, TL_K_CG_W32);
    auto kern = c->gc_fmt == GC_F16 ? k_cgg<4, false, TL_RV_PW32, TL_PIPE_PW32 != 0, GC_F16>
                                    : k_cgg<4, false, TL_RV_PW32, TL_PIPE_PW32 != 0, GC_BF16>;
    kern<<<nb4p(c), TBV, 0, c->stream>>>(
        c->g, c->g.typ4, targ(t, tref), t >= 2 ? p32(c, t - 1) : nullptr, p32(c, t), c->r32, nullptr, c->gcell, c->hx32, c->hy32,
        c->ctrl, nullptr, 0, t >= 2 ? part_r(c, t - 1) : nullptr, t >= 2 ? nbr(c, t - 1) : 0, part_w(c), nullptr);
    c->nb_w = nb4p(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_cg_ur32(tl_ctx *c, int t, int tref)
{
    const bool prof = prof_start(c, TL_K_CG_UR32);
    auto kern = c->gc_fmt == GC_F16 ? k_cgg<4, true, TL_RV_UR32, TL_PIPE_UR32 != 0, GC_F16>
                                    : k_cgg<4, true, TL_RV_UR32, TL_PIPE_UR32 != 0, GC_BF16>;
    kern<<<nb4(c), TBV, 0, c->stream>>>(
        c->g, c->g.ty4, targ(t, tref), nullptr, p32(c, t), c->r32, c->x32, c->gcell, c->hx32, c->hy32, c->ctrl, part_w(c),
        c->nb_w, nullptr, 0, part_r(c, t), t >= 2 ? p32(c, t - 1) : nullptr);
    nbr(c, t) = nb4(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_graph_base(tl_ctx *c, int k)
{
    k_graph_base<<<1, 64, 0, c->stream>>>(c->ctrl, k);
    return hipGetLastError();
}

hipError_t launch_graph_seed(tl_ctx *c, int t0)
{
    k_graph_seed<<<1, 64, 0, c->stream>>>(c->ctrl, t0);
    return hipGetLastError();
}

// the true FP64 residual r = u0 - A u of the FP64 CG after iteration T
// replaces the recursive one (r and its partial sums of iteration T); the
// finalizer then stops on it or lets iteration T+1 run
hipError_t launch_cg64_true_residual(tl_ctx *c, int T)
{
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, 0, c->f[TL_F_U0], c->f[TL_F_U], c->f[TL_F_KX],
                                                  c->f[TL_F_KY], c->f[TL_F_R], nullptr, nullptr, c->ctrl,
                                                  part_r(c, T));
    nbr(c, T) = nbh(c);
    hipError_t e = hipGetLastError();
    // result in sums[3]: sums[0..1] still hold the initial residual and bb
    return e == hipSuccess ? fin(c, part_r(c, T), nbh(c), 1, 3, FK_IR, T) : e;
}

hipError_t launch_residual(tl_ctx *c)
{
    const bool prof = prof_start(c, TL_K_RESID64);
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, 0, c->f[TL_F_U0], c->f[TL_F_U], c->f[TL_F_KX],
                                                  c->f[TL_F_KY], c->f[TL_F_R], nullptr, nullptr, c->ctrl, nullptr);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

hipError_t launch_jacobi_residual(tl_ctx *c, const double *u, int t)
{
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, c->f[TL_F_U0], u, c->f[TL_F_KX], c->f[TL_F_KY],
                                                  c->f[TL_F_R], nullptr, nullptr, c->ctrl, part_m(c) + stride(c));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 2, 0, FK_JACOBI, t) : e;
}

hipError_t launch_ir_residual(tl_ctx *c, int reset_p, int t_next, float *r_dst)
{
    const bool prof = prof_start(c, TL_K_RESID64);
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, 0, c->f[TL_F_U0], c->f[TL_F_U], c->f[TL_F_KX],
                                                  c->f[TL_F_KY], nullptr, r_dst ? r_dst : c->r32,
                                                  reset_p ? c->p32 : nullptr, c->ctrl, part_r(c, t_next - 1));
    if (prof) prof_stop(c);
    nbr(c, t_next - 1) = nbh(c);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_r(c, t_next - 1), nbh(c), 1, 0, FK_IR, t_next - 1) : e;
}

hipError_t launch_apply(tl_ctx *c, const double *src, double *dst)
{
    k_apply<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, src, c->f[TL_F_KX], c->f[TL_F_KY], dst);
    return hipGetLastError();
}

hipError_t launch_finalise(tl_ctx *c)
{
    k_finalise<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], c->f[TL_F_DENSITY], c->f[TL_F_ENERGY]);
    return hipGetLastError();
}

hipError_t launch_field_summary(tl_ctx *c)
{
    k_field_summary<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_DENSITY], c->f[TL_F_ENERGY], c->f[TL_F_U],
                                                       c->dx * c->dy, part_m(c));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 4, 0, FK_SUMS, 0) : e;
}

hipError_t launch_sumsq_slot(tl_ctx *c, const double *v, int slot)
{
    k_sumsq<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, 0, v, c->ctrl, part_m(c));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 1, slot, FK_SUMS, 0) : e;
}

hipError_t launch_sumsq_rr(tl_ctx *c, const double *v, int t)
{
    k_sumsq<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, v, c->ctrl, part_r(c, t));
    nbr(c, t) = nbh(c);
    return hipGetLastError();
}

hipError_t launch_halo(tl_ctx *c, unsigned mask, int depth, double *const *extra, int nextra)
{
    HaloArgs h{};
    h.n = 0;
    static const int order[6] = {TL_F_DENSITY, TL_F_P, TL_F_ENERGY0, TL_F_ENERGY, TL_F_U, TL_F_SD};
    for (int b = 0; b < 6; ++b)
        if (mask & (1u << b)) h.f[h.n++] = c->f[order[b]];
    for (int e = 0; e < nextra && h.n < 8; ++e) h.f[h.n++] = extra[e];
    if (h.n == 0 || depth <= 0) return hipSuccess;
    const int n = c->g.nx > c->g.ny ? c->g.nx : c->g.ny;
    k_halo<<<dim3((n + TB - 1) / TB, h.n, depth), TB, 0, c->stream>>>(c->g, h, depth);
    return hipGetLastError();
}

hipError_t launch_cheby_init(tl_ctx *c, double theta, const double *u_in, double *u_out)
{
    k_cheby_init<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, u_in, u_out, c->f[TL_F_U0], c->f[TL_F_P], c->f[TL_F_R],
                                                    c->f[TL_F_KX], c->f[TL_F_KY], theta, part_m(c));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fin(c, part_m(c), nbh(c), 1, 0, FK_BB, 0) : e;
}

hipError_t launch_cheby_iter(tl_ctx *c, double a, double b, const double *u_in, double *u_out, double *p, int norm,
                             int t)
{
    const bool prof = prof_start(c, TL_K_CHEBY64);
    k_vs64<VS_CHEBY, TL_RV_VS64, TL_PIPE_VS64 != 0><<<nb2(c), TBV, 0, c->stream>>>(
        c->g, c->g.ty2, t, u_in, u_out, c->f[TL_F_KX], c->f[TL_F_KY], c->f[TL_F_U0], p, a, b, norm, c->ctrl,
        part_m(c));
    if (prof) prof_stop(c);
    hipError_t e = hipGetLastError();
    return (e == hipSuccess && norm) ? fin(c, part_m(c), nb2(c), 1, 0, FK_CHEBY, t) : e;
}

// two Chebyshev steps, iterations t and t+1 (coefficients a1, b1 then a2,
// b2): u_in -> u_out, p_in -> p_out; with norm, the residual norm of t+1
hipError_t launch_cheby_iter2(tl_ctx *c, double a1, double b1, double a2, double b2, const double *u_in,
                              double *u_out, const double *p_in, double *p_out, int norm, int t)
{
    const bool prof = prof_start(c, TL_K_CHEBY2_64);
    // the launch that forms the residual norm uses the single step's bands
    const unsigned nb = norm ? nb2(c) : (unsigned)(c->g.gx2 * c->g.gyq);
    k_cheby2<<<nb, TBV, 0, c->stream>>>(c->g, norm ? c->g.ty2 : c->g.tyq, t, u_in, u_out, c->f[TL_F_U0], p_in,
                                        p_out, c->f[TL_F_KX], c->f[TL_F_KY], a1, b1, a2, b2, norm, c->ctrl,
                                        part_m(c));
    if (prof) prof_stop(c);
    hipError_t e = hipGetLastError();
    return (e == hipSuccess && norm) ? fin(c, part_m(c), nb2(c), 1, 0, FK_CHEBY, t + 1) : e;
}

hipError_t launch_ppcg_ur_sd(tl_ctx *c, double theta, int t)
{
    k_ppcg_ur_sd<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, c->f[TL_F_U], c->f[TL_F_P], c->f[TL_F_R],
                                                    c->f[TL_F_W], c->f[TL_F_SD], theta, c->ctrl, part_w(c),
                                                    c->nb_w, part_r(c, t - 1), nbr(c, t - 1));
    return hipGetLastError();
}

hipError_t launch_ppcg_inner(tl_ctx *c, double a, double b, const double *sd_in, double *sd_out, int last, int t)
{
    const bool prof = prof_start(c, TL_K_PPCG64);
    k_vs64<VS_PPCG, TL_RV_VS64, TL_PIPE_VS64 != 0><<<nb2(c), TBV, 0, c->stream>>>(
        c->g, c->g.ty2, t, sd_in, sd_out, c->f[TL_F_KX], c->f[TL_F_KY], c->f[TL_F_R], c->f[TL_F_U], a, b, last,
        c->ctrl, part_r(c, t));
    if (prof) prof_stop(c);
    if (last) nbr(c, t) = nb2(c);
    return hipGetLastError();
}

// two inner steps (coefficients a1, b1 then a2, b2): sd_in -> sd_out,
// r_in -> r_out (distinct buffers), u in place
hipError_t launch_ppcg_inner2(tl_ctx *c, double a1, double b1, double a2, double b2, const double *sd_in,
                              double *sd_out, const double *r_in, double *r_out, int last, int t)
{
    const bool prof = prof_start(c, TL_K_PPCG2_64);
    // the launch that forms the residual sum uses the FP64 kernels' bands
    const unsigned nb = last ? nb2(c) : (unsigned)(c->g.gx2 * c->g.gyq);
    k_ppcg2<<<nb, TBV, 0, c->stream>>>(c->g, last ? c->g.ty2 : c->g.tyq, t, sd_in, sd_out, r_in, r_out, c->f[TL_F_U],
                                           c->f[TL_F_KX], c->f[TL_F_KY], a1, b1, a2, b2, last, c->ctrl,
                                           part_r(c, t));
    if (prof) prof_stop(c);
    if (last) nbr(c, t) = nb2(c);
    return hipGetLastError();
}

hipError_t launch_jacobi(tl_ctx *c, const double *u_in, double *u_out, int t, int defer)
{
    const bool prof = prof_start(c, TL_K_JACOBI64);
    if (defer) {
        // the sweep whose sum joins the residual's (every 50th): the scalar
        // kernel, so that both partial arrays have k_residual's block count
        k_jacobi<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, t, u_in, u_out, c->f[TL_F_U0], c->f[TL_F_KX],
                                                    c->f[TL_F_KY], c->ctrl, part_m(c));
    } else {
        k_vs64<VS_JACOBI, TL_RV_VS64, TL_PIPE_VS64 != 0><<<nb2(c), TBV, 0, c->stream>>>(
            c->g, c->g.ty2, t, u_in, u_out, c->f[TL_F_KX], c->f[TL_F_KY], c->f[TL_F_U0], nullptr, 0.0, 0.0, 1,
            c->ctrl, part_m(c));
    }
    if (prof) prof_stop(c);
    hipError_t e = hipGetLastError();
    return (e == hipSuccess && !defer) ? fin(c, part_m(c), nb2(c), 1, 0, FK_JACOBI, t) : e;
}

// u += d; d = 0 over the whole allocation (d is zero outside the interior)
__global__ __launch_bounds__(TB) void k_acc64(double *__restrict__ u, double *__restrict__ d, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * TB + threadIdx.x; i < n; i += (size_t)gridDim.x * TB) {
        const double v = d[i];
        u[i] += v;
        d[i] = 0.0;
    }
}

hipError_t launch_acc64(tl_ctx *c, double *u, double *d)
{
    const size_t n = c->g.elems;
    const unsigned nb = (unsigned)std::min<size_t>((n + TB - 1) / TB, 8192);
    k_acc64<<<nb, TB, 0, c->stream>>>(u, d, n);
    return hipGetLastError();
}

hipError_t launch_ir_prepare(tl_ctx *c, double rx, double ry, int coefficient)
{
    const size_t n = c->g.elems;
    const unsigned nb = (unsigned)((n + TB - 1) / TB);
    int fmt = GC_BF16;
    float scale = 1.0f;
#if TL_GC16
    // half with a power-of-two scale when the coefficient's range fits:
    // max * scale <= 2^14 and min * scale >= 2^-14 (half's smallest normal)
    k_gc_range<<<TL_RANGE_BLOCKS, TB, 0, c->stream>>>(c->f[TL_F_DENSITY], n, coefficient, c->gc_range);
    hipError_t e = hipGetLastError();
    float h[2 * TL_RANGE_BLOCKS];
    if (e == hipSuccess) e = hipMemcpyAsync(h, c->gc_range, sizeof h, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    float lo = INFINITY, hi = 0.0f;
    for (int b = 0; b < TL_RANGE_BLOCKS; ++b) {
        lo = std::min(lo, h[2 * b]);
        hi = std::max(hi, h[2 * b + 1]);
    }
    if (hi > 0.0f && std::isfinite(hi) && lo > 0.0f) {
        const float s = std::ldexp(1.0f, (int)std::floor(std::log2(16384.0 / (double)hi)));
        if ((double)lo * s >= std::ldexp(1.0, -14)) {
            fmt = GC_F16;
            scale = s;
        }
    }
#endif
    if (fmt == GC_F16)
        k_gcell<GC_F16><<<nb, TB, 0, c->stream>>>(c->f[TL_F_DENSITY], c->gcell, n, coefficient, scale);
    else
        k_gcell<GC_BF16><<<nb, TB, 0, c->stream>>>(c->f[TL_F_DENSITY], c->gcell, n, coefficient, 1.0f);
    c->gc_fmt = fmt;
    c->gc_scale = scale;
    // kx = hx (gc(x-1) + gc(x)): the scale of the stored gc is divided out of hx, hy
    c->hx32 = (float)(0.5 * rx / scale);
    c->hy32 = (float)(0.5 * ry / scale);
    return hipGetLastError();
}

hipError_t launch_ir_accumulate(tl_ctx *c, int T, const float *pT)
{
    k_ir_accumulate<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], c->x32,
                                                       T >= 1 ? (pT ? pT : p32(c, T)) : nullptr, c->ctrl, T);
    return hipGetLastError();
}

// fused FP32 inner CG: r(t) in buffer t & 1, p(t) in buffer t & 1, the four
// per-block sums for iteration t+1 in set t & 1
float *cgf_r(tl_ctx *c, int t) { return (t & 1) ? c->r32b : c->r32; }
float *cgf_p(tl_ctx *c, int t) { return (t & 1) ? c->p32b : c->p32; }
static inline double *part_f(tl_ctx *c, int t) { return c->part_f + (size_t)4 * (t & 1) * stride(c); }
static inline unsigned nbf(const tl_ctx *c) { return (unsigned)(c->g.gx4 * c->g.gyf); }

hipError_t launch_cgf(tl_ctx *c, int t, int tref)
{
    const bool prof = prof_start(c, TL_K_CGF32);
    auto kern = c->gc_fmt == GC_F16 ? k_cgf<TL_RV_F32, GC_F16, TL_CGF_DEPTH> : k_cgf<TL_RV_F32, GC_BF16, TL_CGF_DEPTH>;
    kern<<<nbf(c), TBV, 0, c->stream>>>(c->g, c->g.tyf, targ(t, tref), cgf_r(c, t - 1), cgf_r(c, t),
                                        cgf_p(c, t - 1), cgf_p(c, t), c->x32, c->gcell, c->hx32, c->hy32, c->ctrl,
                                        part_f(c, t - 1), c->nb_f[(t - 1) & 1], part_f(c, t));
    c->nb_f[t & 1] = nbf(c);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

// pairs of FP32 inner CG iterations (k_cgp): after iteration T (even), r(T)
// and p(T) in buffer (T / 2) & 1; the six sums of a pair's first sweep in
// part_f[0 .. 6 stride), the residual sum of its second in part_f[6 stride ..)
float *cgp_r(tl_ctx *c, int T) { return ((T >> 1) & 1) ? c->r32b : c->r32; }
float *cgp_p(tl_ctx *c, int T) { return ((T >> 1) & 1) ? c->p32b : c->p32; }

// the pair t, t+1 (t odd); first: the first pair after a start or an FP64
// refresh, whose gamma(t-1) is in Ctrl::rrs
hipError_t launch_cgp(tl_ctx *c, int t, bool first, int tref)
{
    double *const pg = c->part_f, *const pu = c->part_f + 6 * stride(c);
    const unsigned nb = nbf(c);
    const bool fp16 = c->gc_fmt == GC_F16;
    bool prof = prof_start(c, TL_K_CGP_G32);
    auto kg = fp16 ? k_cgp<GC_F16, false> : k_cgp<GC_BF16, false>;
    kg<<<nb, TBV, 0, c->stream>>>(c->g, c->g.tyf, targ(t, tref), cgp_r(c, t - 1), nullptr, cgp_p(c, t - 1), nullptr,
                                  nullptr, c->gcell, c->hx32, c->hy32, c->ctrl, first ? nullptr : pu, nb, pg);
    if (prof) prof_stop(c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    prof = prof_start(c, TL_K_CGP_U32);
    auto ku = fp16 ? k_cgp<GC_F16, true> : k_cgp<GC_BF16, true>;
    ku<<<nb, TBV, 0, c->stream>>>(c->g, c->g.tyf, targ(t, tref), cgp_r(c, t - 1), cgp_r(c, t + 1), cgp_p(c, t - 1),
                                  cgp_p(c, t + 1), c->x32, c->gcell, c->hx32, c->hy32, c->ctrl, pg, nb, pu);
    if (prof) prof_stop(c);
    return hipGetLastError();
}

// the four sums of r(T) and p(T) for iteration T+1 (after a start or a refresh)
hipError_t launch_cgf_dots(tl_ctx *c, int T)
{
    auto kern = c->gc_fmt == GC_F16 ? k_cgf_dots<GC_F16> : k_cgf_dots<GC_BF16>;
    kern<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, cgf_r(c, T), cgf_p(c, T), c->gcell, c->hx32, c->hy32,
                                            part_f(c, T));
    c->nb_f[T & 1] = nbh(c);
    return hipGetLastError();
}

}  // namespace tl

// ------------------------------------------- field summary without allocation
//
// tl_field_summary_noalloc: the field summary of caller-owned device fields
// in the host layout (X*Y doubles, x fastest), for callers that cannot
// allocate.  Its per-block partial sums live in `scratch`, a static device
// array of the library (one copy per GPU, held until the process ends).
// Host code never passes the array's name to a kernel: it is a host-side
// stand-in there.  The array's device address is taken once per device with
// hipGetSymbolAddress and kept in noalloc_scratch_ptr.

namespace {

constexpr int NOALLOC_MAX_BLOCKS = 8192;  // partial sums of 4 quantities for up to this many blocks
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[4 * NOALLOC_MAX_BLOCKS];

// sum the 4 partial arrays (nb long, stride NOALLOC_MAX_BLOCKS) in the order
// k_fin uses for the context's field summary, into out[0..3]
__global__ __launch_bounds__(tl::TB) void k_sum4_out(const double *__restrict__ part, unsigned nb,
                                                     double *__restrict__ out)
{
    double a[2], b[2];
    tl::sum_partials<2, tl::TB, true>(part, nb, NOALLOC_MAX_BLOCKS, a);
    tl::sum_partials<2, tl::TB, true>(part + 2 * (size_t)NOALLOC_MAX_BLOCKS, nb, NOALLOC_MAX_BLOCKS, b);
    if (threadIdx.x == 0) {
        out[0] = a[0];
        out[1] = a[1];
        out[2] = b[0];
        out[3] = b[1];
    }
}

std::mutex noalloc_mu;
double *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of this device's `scratch` (the current device must be dev)
hipError_t noalloc_scratch_ptr(int dev, double **out)
{
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        noalloc_ptr[dev] = static_cast<double *>(p);
    }
    *out = noalloc_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int tl_field_summary_noalloc(const double *density, const double *energy, const double *u, int nx,
                                        int ny, int halo_depth, double dx, double dy, double *out_dev,
                                        void *stream)
{
    if (!density || !energy || !u || !out_dev || nx < 1 || ny < 1 || halo_depth < 1 || halo_depth > 16)
        return TL_E_INVALID;
    // the layout of the caller's arrays: pitch X, interior (0, 0) at (hd, hd)
    tl::Geo g = tl::make_geo(nx, ny, halo_depth);
    g.pitch = g.X;
    g.off = (long)halo_depth * g.X + halo_depth;
    const long nb = (long)g.gx * g.gy;
    if (nb > NOALLOC_MAX_BLOCKS) return TL_E_INVALID;  // beyond the static partial-sum array
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return TL_E_HIP;
    double *part = nullptr;
    if (noalloc_scratch_ptr(dev, &part) != hipSuccess) return TL_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // one partial per block and quantity at part[q * NOALLOC_MAX_BLOCKS + block]
    g.nbmax = NOALLOC_MAX_BLOCKS;
    tl::k_field_summary<<<dim3(g.gx, g.gy), tl::TB, 0, s>>>(g, density, energy, u, dx * dy, part);
    if (hipGetLastError() != hipSuccess) return TL_E_HIP;
    k_sum4_out<<<1, tl::TB, 0, s>>>(part, (unsigned)nb, out_dev);
    return hipGetLastError() == hipSuccess ? TL_OK : TL_E_HIP;
}
