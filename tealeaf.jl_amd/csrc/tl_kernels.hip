// tl_kernels.hip -- gfx950 kernels of the TeaLeaf linear-solve hot path.
//
// All stencil kernels share one launch shape: a block of 256 threads owns a
// 256-cell row segment and sweeps ty consecutive rows, keeping the cells
// above and below in registers (a vertical sliding window) so every field is
// fetched from HBM once per sweep.  Left/right neighbours are re-read through
// L1/L2.  The depth-1 reflective halo of kernels.jl:191-210 (with the index
// correction of the intended algorithm) is applied by clamping neighbour
// indices to the interior: a reflected ghost cell holds exactly the value of
// the interior cell it mirrors, so the arithmetic is bit-identical to
// reading an updated halo, and no halo exchange is needed inside the solve.
//
// Reductions are single-pass and deterministic: wave64 butterfly sums, a
// fixed-order block sum, one partial per block, and the last block (atomic
// ticket) sums the partials in block order and runs the scalar update of the
// solver (alpha, beta, stop test) on the device.
#include <cmath>

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
    long rows = (long)ny * g.gx;
    int ty = (int)((rows + 4095) / 4096);
    if (ty < 2) ty = 2;
    if (ty > 128) ty = 128;
    g.ty = ty;
    g.gy = (ny + ty - 1) / ty;
    return g;
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums v[] over the whole grid.  Returns true in the last block to finish,
// where every thread then holds the grid totals in tot[].
template <int N>
__device__ __forceinline__ bool grid_reduce(const double (&v)[N], double *__restrict__ part, unsigned *ticket,
                                            double (&tot)[N])
{
    __shared__ double sh[N][TB / 64];
    __shared__ int am_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        const double s = wave_sum(v[a]);
        if (lane == 0) sh[a][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < N; ++a) {
            double s = sh[a][0];
#pragma unroll
            for (int q = 1; q < TB / 64; ++q) s += sh[a][q];
            part[(size_t)a * nb + bid] = s;
        }
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        am_last = (t == nb - 1);
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
#pragma unroll
    for (int a = 0; a < N; ++a) {
        double s = 0.0;
        for (unsigned q = threadIdx.x; q < nb; q += TB) s += part[(size_t)a * nb + q];
        s = wave_sum(s);
        __syncthreads();
        if (lane == 0) sh[a][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < N; ++a) {
        double s = sh[a][0];
#pragma unroll
        for (int q = 1; q < TB / 64; ++q) s += sh[a][q];
        tot[a] = s;
    }
    if (threadIdx.x == 0) *ticket = 0u;
    return true;
}

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

// Iterates the rows of this thread's column: calls body(i, y) with i the
// element index of interior cell (x, y).  xl/xr are the clamped horizontal
// neighbour offsets.
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

// Vertical sliding window over field a: cur/dn/up values with the
// reflective clamp at y = 0 and y = ny-1.
template <typename T>
struct Window {
    T dn, cur;
    __device__ __forceinline__ void start(const T *__restrict__ a, long i, int y0, long pitch)
    {
        cur = a[i];
        dn = y0 > 0 ? a[i - pitch] : cur;
    }
    __device__ __forceinline__ T up(const T *__restrict__ a, long i, int y, int ny, long pitch) const
    {
        return y + 1 < ny ? a[i + pitch] : cur;
    }
};

// ------------------------------------------------------- scalar updates

__device__ void decide(Ctrl *c, double rr)
{
    const int it = c->iters;
    const double nr = sqrt(fabs(rr));
    int why = WHY_NONE;
    switch (c->mode) {
    case MODE_CG_REF:
        if (nr < c->eps) why = WHY_CONVERGED;
        break;
    case MODE_CG_REL:
        if (nr <= c->tol_abs) why = WHY_CONVERGED;
        break;
    case MODE_SWITCH: {
        if (fabs(rr) < c->eps) {
            why = WHY_CONVERGED;
        } else {
            const int tt = it + 1;
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
    if (why == WHY_NONE && it >= c->maxiters) why = WHY_MAXITERS;
    if (!isfinite(rr)) {
        why = WHY_BREAKDOWN;
        c->status = TL_E_BREAKDOWN;
    }
    if (why != WHY_NONE) {
        c->why = why;
        c->F = 1;
    }
}

// -------------------------------------------------------- init kernels

__global__ __launch_bounds__(TB) void k_init_ctrl(Ctrl *c, int mode, int maxiters, double eps, double rel_tol,
                                                  double ir_delta, int presteps, int errorswitch, double epslim)
{
    if (threadIdx.x != 0) return;
    c->rro = c->pw = c->alpha = c->beta = 0.0;
    c->err = 1e10; /* ERROR_START, kernels.jl:7 */
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
    c->F = 0;
    c->G = 0;
    c->why = WHY_NONE;
    c->status = 0;
    c->ticket = 0u;
}

// switches the stop rule between phases; clears the stop flags
__global__ __launch_bounds__(TB) void k_set_mode(Ctrl *c, int mode, int maxiters)
{
    if (threadIdx.x != 0) return;
    c->mode = mode;
    c->maxiters = maxiters;
    c->iters = 0;
    c->F = 0;
    c->G = 0;
    c->why = WHY_NONE;
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

// CG.init! lines 70-78 + copyu! (kernels.jl:217-220), with an optional
// initial guess: b = u (energy*density); x = guess or b;
// w = A x; r = b - w; p = r; u0 = b; u = x.   rro = sum r^2, bb = sum b^2.
__global__ __launch_bounds__(TB) void k_cg_init_resid(Geo g, double *__restrict__ u,
                                                      const double *__restrict__ guess, double *__restrict__ u0,
                                                      double *__restrict__ p, double *__restrict__ r,
                                                      double *__restrict__ w, const double *__restrict__ kx,
                                                      const double *__restrict__ ky, Ctrl *__restrict__ ctrl,
                                                      double *__restrict__ part)
{
    const Col c = column(g);
    double acc[2] = {0.0, 0.0};
    const double *xs = guess ? guess : u;
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(xs, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double xu = win.up(xs, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double b = u[i];
            const double wv = smvp5(win.cur, xs[i + c.xl], xs[i + c.xr], win.dn, xu, kx[i], kx[i + 1], kyd, kyu);
            const double rv = b - wv;
            w[i] = wv;
            r[i] = rv;
            p[i] = rv;
            u0[i] = b;
            if (guess) u[i] = win.cur;
            acc[0] += rv * rv;
            acc[1] += b * b;
            win.dn = win.cur;
            win.cur = xu;
            kyd = kyu;
        }
    }
    double tot[2];
    if (grid_reduce<2>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        ctrl->rro = tot[0];
        ctrl->err = tot[0];
        ctrl->bb = tot[1];
        ctrl->tol_abs = ctrl->rel_tol * sqrt(tot[1]);
        ctrl->refresh_abs = ctrl->ir_delta * sqrt(tot[0]);
        ctrl->sums[0] = tot[0];
        ctrl->sums[1] = tot[1];
        if (tot[0] == 0.0) {  // exact solution already: nothing to iterate
            ctrl->F = 1;
            ctrl->G = 1;
            ctrl->why = WHY_CONVERGED;
        }
    }
}

// sum over the interior of (b - A x)^2 with b = u, into ctrl->sums[slot]
__global__ __launch_bounds__(TB) void k_resid_norm(Geo g, const double *__restrict__ b,
                                                   const double *__restrict__ xs, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, Ctrl *__restrict__ ctrl,
                                                   double *__restrict__ part, int slot)
{
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(xs, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double xu = win.up(xs, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double rv = b[i] - smvp5(win.cur, xs[i + c.xl], xs[i + c.xr], win.dn, xu, kx[i], kx[i + 1], kyd, kyu);
            acc[0] += rv * rv;
            win.dn = win.cur;
            win.cur = xu;
            kyd = kyu;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) ctrl->sums[slot] = tot[0];
}

// ------------------------------------------------------------ CG kernels

// CG.w!, CG.jl:82-90: w = A p; pw = sum p*w; then alpha = rro/pw (CG.jl:35-36)
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_w(Geo g, const T *__restrict__ p, const T *__restrict__ kx,
                                             const T *__restrict__ ky, T *__restrict__ w, Ctrl *__restrict__ ctrl,
                                             double *__restrict__ part, double *__restrict__ hist_a)
{
    if (ctrl->F) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) ctrl->G = 1;
        return;
    }
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<T> win;
        win.start(p, i, c.y0, g.pitch);
        T kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const T pu = win.up(p, i, y, g.ny, g.pitch);
            const T kyu = ky[i + g.pitch];
            const T wv = smvp5(win.cur, p[i + c.xl], p[i + c.xr], win.dn, pu, kx[i], kx[i + 1], kyd, kyu);
            w[i] = wv;
            acc[0] += (double)wv * (double)win.cur;
            win.dn = win.cur;
            win.cur = pu;
            kyd = kyu;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        const double pw = tot[0];
        const double alpha = ctrl->rro / pw;
        ctrl->pw = pw;
        ctrl->alpha = alpha;
        if (hist_a) hist_a[ctrl->iters] = alpha;
        if (!(pw > 0.0) || !isfinite(alpha)) {
            ctrl->status = TL_E_BREAKDOWN;
            ctrl->why = WHY_BREAKDOWN;
            ctrl->F = 1;
            ctrl->G = 1;
        }
    }
}

// CG.ur!, CG.jl:93-98: u += alpha p; r -= alpha w; rr = sum r^2; then
// beta = rrn/rro (CG.jl:39-40) and the stop test
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_ur(Geo g, T *__restrict__ u, const T *__restrict__ p,
                                              T *__restrict__ r, const T *__restrict__ w, Ctrl *__restrict__ ctrl,
                                              double *__restrict__ part, double *__restrict__ hist_b)
{
    if (ctrl->F) return;
    const T alpha = (T)ctrl->alpha;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            u[i] = u[i] + alpha * p[i];
            const T rv = r[i] - alpha * w[i];
            r[i] = rv;
            acc[0] += (double)rv * (double)rv;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        const double rrn = tot[0];
        const double beta = rrn / ctrl->rro;
        ctrl->beta = beta;
        if (hist_b) hist_b[ctrl->iters] = beta;
        ctrl->rro = rrn;
        ctrl->err = rrn;
        ctrl->iters += 1;
        decide(ctrl, rrn);
    }
}

// CG.p!, CG.jl:101-104: p = beta p + r
template <typename T>
__global__ __launch_bounds__(TB) void k_cg_p(Geo g, T *__restrict__ p, const T *__restrict__ r,
                                             const Ctrl *__restrict__ ctrl)
{
    if (ctrl->G) return;
    const T beta = (T)ctrl->beta;
    const Col c = column(g);
    if (!c.active) return;
    long i = g.off + (long)c.y0 * g.pitch + c.x;
    for (int y = c.y0; y < c.y1; ++y, i += g.pitch) p[i] = beta * p[i] + r[i];
}

// --------------------------------------------------- residual / finalise

// residual!, kernels.jl:227-232: r = u0 - A u; optional sum r^2 with a finalizer
__global__ __launch_bounds__(TB) void k_residual(Geo g, const double *__restrict__ u0, const double *__restrict__ u,
                                                 const double *__restrict__ kx, const double *__restrict__ ky,
                                                 double *__restrict__ r, float *__restrict__ r32,
                                                 float *__restrict__ p32, Ctrl *__restrict__ ctrl,
                                                 double *__restrict__ part, int with_norm, int fin)
{
    if (fin == FIN_JACOBI_ADD && ctrl->F) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(u, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double uu = win.up(u, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double rv = u0[i] - smvp5(win.cur, u[i + c.xl], u[i + c.xr], win.dn, uu, kx[i], kx[i + 1], kyd, kyu);
            if (r) r[i] = rv;
            if (r32) {
                const float rf = (float)rv;
                r32[i] = rf;
                if (p32) p32[i] = rf;
            }
            acc[0] += rv * rv;
            win.dn = win.cur;
            win.cur = uu;
            kyd = kyu;
        }
    }
    if (!with_norm) return;
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        const double rr = tot[0];
        ctrl->sums[0] = rr;
        if (fin == FIN_JACOBI_ADD) {
            ctrl->err += rr;  // Jacobi.jl:19-20
            decide(ctrl, ctrl->err);
        } else if (fin == FIN_IR) {
            // FP64 refresh of the mixed-precision CG: the true residual
            // replaces the FP32 recurrence's; converged or resume
            ctrl->rro = rr;
            ctrl->err = rr;
            ctrl->refresh_abs = ctrl->ir_delta * sqrt(rr);
            ctrl->G = 0;
            if (sqrt(rr) <= ctrl->tol_abs) {
                ctrl->why = WHY_CONVERGED;
                ctrl->F = 1;
            } else if (ctrl->iters >= ctrl->maxiters || !isfinite(rr)) {
                ctrl->why = isfinite(rr) ? WHY_MAXITERS : WHY_BREAKDOWN;
                if (!isfinite(rr)) ctrl->status = TL_E_BREAKDOWN;
                ctrl->F = 1;
            } else {
                ctrl->why = WHY_NONE;
                ctrl->F = 0;
            }
        }
    }
}

// dst = A src over the interior
__global__ __launch_bounds__(TB) void k_apply(Geo g, const double *__restrict__ a, const double *__restrict__ kx,
                                              const double *__restrict__ ky, double *__restrict__ o)
{
    const Col c = column(g);
    if (!c.active) return;
    long i = g.off + (long)c.y0 * g.pitch + c.x;
    Window<double> win;
    win.start(a, i, c.y0, g.pitch);
    double kyd = ky[i];
    for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
        const double au = win.up(a, i, y, g.ny, g.pitch);
        const double kyu = ky[i + g.pitch];
        o[i] = smvp5(win.cur, a[i + c.xl], a[i + c.xr], win.dn, au, kx[i], kx[i + 1], kyd, kyu);
        win.dn = win.cur;
        win.cur = au;
        kyd = kyu;
    }
}

// finalise!, kernels.jl:239-242: energy = u / density
__global__ __launch_bounds__(TB) void k_finalise(Geo g, const double *__restrict__ u, const double *__restrict__ rho,
                                                 double *__restrict__ e)
{
    const Col c = column(g);
    if (!c.active) return;
    long i = g.off + (long)c.y0 * g.pitch + c.x;
    for (int y = c.y0; y < c.y1; ++y, i += g.pitch) e[i] = u[i] / rho[i];
}

// fieldsummary, kernels.jl:119-133: vol, mass, ie, temp
__global__ __launch_bounds__(TB) void k_field_summary(Geo g, const double *__restrict__ rho,
                                                      const double *__restrict__ e, const double *__restrict__ u,
                                                      double vol, Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    const Col c = column(g);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double m = vol * rho[i];
            acc[0] += vol;
            acc[1] += m;
            acc[2] += m * e[i];
            acc[3] += m * u[i];
        }
    }
    double tot[4];
    if (grid_reduce<4>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0)
        for (int a = 0; a < 4; ++a) ctrl->sums[a] = tot[a];
}

// sum v^2 over the interior with a finalizer (PPCG norm when there are no inner steps)
__global__ __launch_bounds__(TB) void k_sumsq(Geo g, const double *__restrict__ v, Ctrl *__restrict__ ctrl,
                                              double *__restrict__ part, int fin)
{
    if (fin == FIN_PPCG && ctrl->F) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) acc[0] += v[i] * v[i];
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        ctrl->sums[0] = tot[0];
        if (fin == FIN_PPCG) {
            const double rrn = tot[0];
            ctrl->beta = rrn / ctrl->rro;
            ctrl->rro = rrn;
            ctrl->err = rrn;
            ctrl->iters += 1;
            decide(ctrl, rrn);
        }
    }
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
// bb = sum u0^2 (u' is a separate buffer: the stencil reads u's neighbours)
__global__ __launch_bounds__(TB) void k_cheby_init(Geo g, const double *__restrict__ uin, double *__restrict__ uout,
                                                   const double *__restrict__ u0, double *__restrict__ p,
                                                   double *__restrict__ r, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double theta,
                                                   Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(uin, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double uu = win.up(uin, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double b = u0[i];
            const double wv = smvp5(win.cur, uin[i + c.xl], uin[i + c.xr], win.dn, uu, kx[i], kx[i + 1], kyd, kyu);
            const double rv = b - wv;
            const double pv = rv / theta;
            r[i] = rv;
            p[i] = pv;
            uout[i] = win.cur + pv;
            acc[0] += b * b;
            win.dn = win.cur;
            win.cur = uu;
            kyd = kyu;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) ctrl->bb = tot[0];
}

// Cheby.mainstep!, Cheby.jl:84-106: w = A u; r = u0 - w; p = a p + b r;
// u' = u + p; optional err = sum r^2 (calc2norm) with the stop test.
// `it` is the Chebyshev iteration index, recorded when this launch decides.
__global__ __launch_bounds__(TB) void k_cheby_iter(Geo g, const double *__restrict__ uin, double *__restrict__ uout,
                                                   const double *__restrict__ u0, double *__restrict__ p,
                                                   const double *__restrict__ kx, const double *__restrict__ ky,
                                                   double ca, double cb, int norm, int it,
                                                   Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (ctrl->F) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(uin, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double uu = win.up(uin, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double wv = smvp5(win.cur, uin[i + c.xl], uin[i + c.xr], win.dn, uu, kx[i], kx[i + 1], kyd, kyu);
            const double rv = u0[i] - wv;
            const double pv = ca * p[i] + cb * rv;
            p[i] = pv;
            uout[i] = win.cur + pv;
            acc[0] += rv * rv;
            win.dn = win.cur;
            win.cur = uu;
            kyd = kyu;
        }
    }
    if (!norm) return;
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        ctrl->err = tot[0];
        ctrl->iters = it;
        if (fabs(tot[0]) < ctrl->eps) {  // Cheby.jl:57
            ctrl->why = WHY_CONVERGED;
            ctrl->F = 1;
        }
    }
}

// ----------------------------------------------------------------- PPCG

// CG.ur! without its norm (PPCG.jl:67, value unused) fused with
// PPCG.init!(chunk, hd) (PPCG.jl:97-100): u += alpha p; r -= alpha w; sd = r/theta
__global__ __launch_bounds__(TB) void k_ppcg_ur_sd(Geo g, double *__restrict__ u, const double *__restrict__ p,
                                                   double *__restrict__ r, const double *__restrict__ w,
                                                   double *__restrict__ sd, double theta,
                                                   const Ctrl *__restrict__ ctrl)
{
    if (ctrl->F) return;
    const double alpha = ctrl->alpha;
    const Col c = column(g);
    if (!c.active) return;
    long i = g.off + (long)c.y0 * g.pitch + c.x;
    for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
        u[i] = u[i] + alpha * p[i];
        const double rv = r[i] - alpha * w[i];
        r[i] = rv;
        sd[i] = rv / theta;
    }
}

// PPCG inner step (PPCG.jl:75-84 in the two-phase form): r -= A sd;
// u += sd; sd' = a sd + b r.  On the last step: rrn = sum r^2 (PPCG.jl:88),
// beta = rrn/rro and the stop test.
__global__ __launch_bounds__(TB) void k_ppcg_inner(Geo g, const double *__restrict__ sdin,
                                                   double *__restrict__ sdout, double *__restrict__ r,
                                                   double *__restrict__ u, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double ca, double cb, int last,
                                                   Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (ctrl->F) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(sdin, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double su = win.up(sdin, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double s = win.cur;
            const double rv = r[i] - smvp5(s, sdin[i + c.xl], sdin[i + c.xr], win.dn, su, kx[i], kx[i + 1], kyd, kyu);
            r[i] = rv;
            u[i] = u[i] + s;
            sdout[i] = ca * s + cb * rv;
            acc[0] += rv * rv;
            win.dn = s;
            win.cur = su;
            kyd = kyu;
        }
    }
    if (!last) return;
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        const double rrn = tot[0];
        ctrl->beta = rrn / ctrl->rro;
        ctrl->rro = rrn;
        ctrl->err = rrn;
        ctrl->iters += 1;
        decide(ctrl, rrn);
    }
}

// --------------------------------------------------------------- Jacobi

// Jacobi.iterate!, Jacobi.jl:59-80 (u' in a separate buffer, the reference's
// r = copy of u): u' = (u0 + sum K u_nbr) / (1 + sum K); err = sum |u' - u|.
// `defer`: the residual norm of Jacobi.jl:16-21 follows and decides.
__global__ __launch_bounds__(TB) void k_jacobi(Geo g, const double *__restrict__ uin, double *__restrict__ uout,
                                               const double *__restrict__ u0, const double *__restrict__ kx,
                                               const double *__restrict__ ky, int tt, int defer,
                                               Ctrl *__restrict__ ctrl, double *__restrict__ part)
{
    if (ctrl->F) return;
    const Col c = column(g);
    double acc[1] = {0.0};
    if (c.active) {
        long i = g.off + (long)c.y0 * g.pitch + c.x;
        Window<double> win;
        win.start(uin, i, c.y0, g.pitch);
        double kyd = ky[i];
        for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
            const double uu = win.up(uin, i, y, g.ny, g.pitch);
            const double kyu = ky[i + g.pitch];
            const double kxl = kx[i], kxr = kx[i + 1];
            double num = u0[i] + kxr * uin[i + c.xr];
            num = num + kxl * uin[i + c.xl];
            num = num + kyu * uu;
            num = num + kyd * win.dn;
            double den = 1.0 + kxl;
            den = den + kxr;
            den = den + kyd;
            den = den + kyu;
            const double un = num / den;
            uout[i] = un;
            acc[0] += fabs(un - win.cur);
            win.dn = win.cur;
            win.cur = uu;
            kyd = kyu;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &ctrl->ticket, tot) && threadIdx.x == 0) {
        ctrl->err = tot[0];
        ctrl->iters = tt;
        if (!defer) decide(ctrl, tot[0]);
    }
}

// ------------------------------------------------------ mixed precision

// FP32 copies of the face coefficients over the whole allocation
__global__ __launch_bounds__(TB) void k_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// FP64 solution accumulation of the mixed-precision CG: u += x; x = 0
__global__ __launch_bounds__(TB) void k_ir_accumulate(Geo g, double *__restrict__ u, float *__restrict__ x)
{
    const Col c = column(g);
    if (!c.active) return;
    long i = g.off + (long)c.y0 * g.pitch + c.x;
    for (int y = c.y0; y < c.y1; ++y, i += g.pitch) {
        u[i] = u[i] + (double)x[i];
        x[i] = 0.0f;
    }
}

// ------------------------------------------------------------- launchers

static inline dim3 sgrid(const Geo &g) { return dim3(g.gx, g.gy); }
static inline dim3 fgrid(const Geo &g) { return dim3((g.X + TB - 1) / TB, g.Y); }

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

hipError_t launch_copy_energy(tl_ctx *c)
{
    k_copy_full<<<fgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_ENERGY0], c->f[TL_F_ENERGY]);
    return hipGetLastError();
}

hipError_t launch_copy_full(tl_ctx *c, const double *a, double *b)
{
    k_copy_full<<<fgrid(c->g), TB, 0, c->stream>>>(c->g, a, b);
    return hipGetLastError();
}

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
                                                       c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_resid_norm_slot(tl_ctx *c, const double *b, const double *x, int slot)
{
    k_resid_norm<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, b, x, c->f[TL_F_KX], c->f[TL_F_KY], c->ctrl, c->partials,
                                                    slot);
    return hipGetLastError();
}

hipError_t launch_cg_w64(tl_ctx *c)
{
    k_cg_w<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_P], c->f[TL_F_KX], c->f[TL_F_KY], c->f[TL_F_W],
                                                       c->ctrl, c->partials, c->hist_a);
    return hipGetLastError();
}

hipError_t launch_cg_ur64(tl_ctx *c)
{
    k_cg_ur<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], c->f[TL_F_P], c->f[TL_F_R], c->f[TL_F_W],
                                                        c->ctrl, c->partials, c->hist_b);
    return hipGetLastError();
}

hipError_t launch_cg_p64(tl_ctx *c)
{
    k_cg_p<double><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_P], c->f[TL_F_R], c->ctrl);
    return hipGetLastError();
}

hipError_t launch_cg_w32(tl_ctx *c)
{
    k_cg_w<float><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->p32, c->kx32, c->ky32, c->w32, c->ctrl, c->partials,
                                                      nullptr);
    return hipGetLastError();
}

hipError_t launch_cg_ur32(tl_ctx *c)
{
    k_cg_ur<float><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->x32, c->p32, c->r32, c->w32, c->ctrl, c->partials,
                                                       nullptr);
    return hipGetLastError();
}

hipError_t launch_cg_p32(tl_ctx *c)
{
    k_cg_p<float><<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->p32, c->r32, c->ctrl);
    return hipGetLastError();
}

hipError_t launch_residual(tl_ctx *c, int with_norm, int finalize_kind)
{
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U0], c->f[TL_F_U], c->f[TL_F_KX], c->f[TL_F_KY],
                                                  c->f[TL_F_R], nullptr, nullptr, c->ctrl, c->partials, with_norm,
                                                  finalize_kind);
    return hipGetLastError();
}

hipError_t launch_residual_u(tl_ctx *c, const double *u, int with_norm, int finalize_kind)
{
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U0], u, c->f[TL_F_KX], c->f[TL_F_KY],
                                                  c->f[TL_F_R], nullptr, nullptr, c->ctrl, c->partials, with_norm,
                                                  finalize_kind);
    return hipGetLastError();
}

hipError_t launch_ir_residual(tl_ctx *c, int reset_p)
{
    k_residual<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U0], c->f[TL_F_U], c->f[TL_F_KX], c->f[TL_F_KY],
                                                  nullptr, c->r32, reset_p ? c->p32 : nullptr, c->ctrl, c->partials,
                                                  1, FIN_IR);
    return hipGetLastError();
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
                                                       c->dx * c->dy, c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_sumsq(tl_ctx *c, const double *v, int finalize_kind)
{
    k_sumsq<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, v, c->ctrl, c->partials, finalize_kind);
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
                                                    c->f[TL_F_KX], c->f[TL_F_KY], theta, c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_cheby_iter2(tl_ctx *c, double a, double b, const double *u_in, double *u_out, int norm, int it)
{
    k_cheby_iter<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, u_in, u_out, c->f[TL_F_U0], c->f[TL_F_P], c->f[TL_F_KX],
                                                    c->f[TL_F_KY], a, b, norm, it, c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_ppcg_ur_sd(tl_ctx *c, double theta)
{
    k_ppcg_ur_sd<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], c->f[TL_F_P], c->f[TL_F_R], c->f[TL_F_W],
                                                    c->f[TL_F_SD], theta, c->ctrl);
    return hipGetLastError();
}

hipError_t launch_ppcg_inner(tl_ctx *c, double a, double b, const double *sd_in, double *sd_out, int last)
{
    k_ppcg_inner<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, sd_in, sd_out, c->f[TL_F_R], c->f[TL_F_U], c->f[TL_F_KX],
                                                    c->f[TL_F_KY], a, b, last, c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_jacobi2(tl_ctx *c, const double *u_in, double *u_out, int tt, int defer)
{
    k_jacobi<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, u_in, u_out, c->f[TL_F_U0], c->f[TL_F_KX], c->f[TL_F_KY], tt,
                                                defer, c->ctrl, c->partials);
    return hipGetLastError();
}

hipError_t launch_ir_prepare(tl_ctx *c)
{
    const size_t n = c->g.elems;
    const unsigned nb = (unsigned)((n + TB - 1) / TB);
    k_to_f32<<<nb, TB, 0, c->stream>>>(c->f[TL_F_KX], c->kx32, n);
    k_to_f32<<<nb, TB, 0, c->stream>>>(c->f[TL_F_KY], c->ky32, n);
    return hipGetLastError();
}

hipError_t launch_ir_accumulate(tl_ctx *c)
{
    k_ir_accumulate<<<sgrid(c->g), TB, 0, c->stream>>>(c->g, c->f[TL_F_U], c->x32);
    return hipGetLastError();
}

}  // namespace tl
