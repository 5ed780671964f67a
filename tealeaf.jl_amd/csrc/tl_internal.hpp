// tl_internal.hpp -- shared definitions of libtealeaf_hip (not part of the ABI).
//
// Device layout.  Each field is one pitched allocation.  The reference's
// 0-based element (k, j) of an X*Y Chunk array (chunk.jl:19-38) lives at
//     j * pitch + k + (32 - hd)
// so the first interior cell of every row (k = hd) sits at column 32, i.e.
// 256-byte aligned for FP64 and 128-byte aligned for FP32, and
// pitch = roundup(nx + hd + 32, 64) elements.  Interior coordinates (x, y),
// 0 <= x < nx, 0 <= y < ny, address  off + y * pitch + x  with
// off = hd * pitch + 32.  The face coefficient kx(x, y) couples cells x-1 and
// x (the reference's kx[k] with k = x + hd); ky likewise in y.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tealeaf_hip.h"

namespace tl {

constexpr int TB = 256;        // threads per block (one row segment of 256 cells)
constexpr int XCOL = 32;       // storage column of the first interior cell
// vectorised CG kernels: threads per block, rows per load batch, and the
// block count their geometry aims at on large grids
#ifndef TL_TBV
#define TL_TBV 256
#endif
constexpr int TBV = TL_TBV;
#ifndef TL_RV
#define TL_RV 4
#endif
// rows per load batch of the FP32 inner-CG kernels (PW, UR), and whether the
// next batch's loads are issued before the current batch is computed (two
// register sets); measured at 8192^2 in one process: PW 2 rows pipelined 3%
// faster than 4 rows plain, UR 4 rows pipelined 3.5% (even iterations) and
// 12% (odd) faster than 8 rows plain
#ifndef TL_RV_PW32
#define TL_RV_PW32 2
#endif
#ifndef TL_RV_UR32
#define TL_RV_UR32 4
#endif
#ifndef TL_PIPE_PW32
#define TL_PIPE_PW32 1
#endif
#ifndef TL_PIPE_UR32
#define TL_PIPE_UR32 1
#endif
// the same for the FP64 CG kernels (measured: PW 4 rows pipelined 2.6%
// faster than plain; the streaming UR gains nothing from it)
#ifndef TL_RV_PW64
#define TL_RV_PW64 TL_RV
#endif
#ifndef TL_RV_UR64
#define TL_RV_UR64 TL_RV
#endif
#ifndef TL_PIPE_PW64
#define TL_PIPE_PW64 1
#endif
#ifndef TL_PIPE_UR64
#define TL_PIPE_UR64 0
#endif
#ifndef TL_VEC_BLOCKS32
#define TL_VEC_BLOCKS32 2048
#endif
#ifndef TL_VEC_BLOCKS32_PW
#define TL_VEC_BLOCKS32_PW 1024
#endif
#ifndef TL_VEC_BLOCKS64
#define TL_VEC_BLOCKS64 8192
#endif
// rows per load batch of the vectorised FP64 Chebyshev / PPCG / Jacobi
// kernels, and whether the next batch's loads are issued before the current
// one is computed.  Measured at 8192^2 (A/B in one box): 4 rows pipelined
// (236 VGPRs, 2 waves per SIMD) against 2 rows: Chebyshev 5.28 / 5.18 TB/s,
// PPCG 5.18 / 5.11, Jacobi 5.39 / 5.11; 4 rows unpipelined, 3 rows, and
// 12288 or 16384 blocks no better
#ifndef TL_RV_VS64
#define TL_RV_VS64 4
#endif
#ifndef TL_PIPE_VS64
#define TL_PIPE_VS64 1
#endif
// the fused FP32 inner-CG kernel (k_cgf, software-pipelined): rows per load
// batch, block count on large grids and the least rows per band (each band
// re-reads 2 rows below and 2 above it).  Measured at 8192^2 (A/B runs
// on one box): 1-row batches pipelined (143 VGPRs, 3 waves per SIMD) in 768
// blocks -- one block per 3 waves of every SIMD, a single round -- take
// 0.314 ms per iteration; 2-row batches (188 VGPRs) 0.317 ms with 1024
// blocks, 0.352 ms with 768; 4-row batches 0.36 ms.  With non-temporal
// sweeps: 768 blocks 0.287 ms, 1024 0.35, 1152 0.36-0.38, 1536 0.295;
// 4 waves per SIMD forced (TL_WPE_F32=4: 128 VGPRs, 36 spilled) 0.37-0.38
// The wave-edge loads (lanes 0 / 63: r, p and the coefficients of the two
// columns outside the wave) cost about 4% of the iteration; one 8-byte load
// by lanes 0-5 passed on by readlane instead of three ran no faster.
#ifndef TL_RV_F32
#define TL_RV_F32 1
#endif
#ifndef TL_VEC_BLOCKS_F32
#define TL_VEC_BLOCKS_F32 768
#endif
#ifndef TL_MIN_ROWS_F32
#define TL_MIN_ROWS_F32 32
#endif
// rows of the fused FP32 kernel staged through the LDS per wave (0: loads
// into registers, one row ahead).  Measured at 8192^2 (A/B in one box):
// depth 3 (116 VGPRs, 44 KB LDS per block, 3 waves per SIMD) 0.312 ms per
// iteration against 0.311 ms for register loads; depth 2 (5 waves per SIMD)
// and 4 (2 waves) 2-5% slower.  A band-walking copy with the same arrays and
// geometry streams at 5.0 TB/s, so more rows in flight have little to win.
// blocks of the two-step PPCG kernel on large grids (its launches that do
// not form the residual sum; the one that does keeps the FP64 kernels'
// bands, so that the sum is the single step's bit for bit)
#ifndef TL_VEC_BLOCKS_PPCG2
#define TL_VEC_BLOCKS_PPCG2 768
#endif
#ifndef TL_CGF_DEPTH
#define TL_CGF_DEPTH 0
#endif
// TL_FP32_KERNELS_AUTO takes the pairs of iterations (k_cgp) where the fused
// kernel fits and launches are stream launches (1), or the fused kernel (0).
// Measured at 8192^2: 0.245 ms per iteration (first sweep 0.145, second
// 0.345 per pair) against 0.287 for the fused kernel, at the same iteration
// counts.
// the pair's second step: A p1 by the stencil of p1 formed on three rows (1),
// or as A r - alpha0 A^2 p + beta0 A p (0)
#ifndef TL_CGP_AP1
#define TL_CGP_AP1 1
#endif
#ifndef TL_CGP_AUTO
#define TL_CGP_AUTO 1
#endif
// neighbour exchange of the FP32 PW kernel by DPP wave shifts (else LDS shuffles)
#ifndef TL_DPP
#define TL_DPP 1
#endif
// FP32 inner CG: lanes 0/63 load the neighbour column with the batched row
// loads instead of when the row is used
#ifndef TL_EDGE_PF
#define TL_EDGE_PF 1
#endif
// block-to-tile mapping of the vectorised kernels (blocks go round-robin to
// the 8 XCDs, block b to XCD b % 8): 0 row-major; 1 each XCD a contiguous
// slab of bands; 2 groups of 8 bands, band 8q + x on XCD x with its tiles in
// consecutive slots, so horizontally adjacent tiles share one L2.  Measured
// at 8192^2 in A/B runs on one box: 2 makes the FP32 PW 2%, the FP32 UR 4%
// (odd t) and the FP64 PW 1.4% faster than 0; 1 is 2-4% slower than 0.
#ifndef TL_XCD_MAP
#define TL_XCD_MAP 2
#endif
// optional occupancy target (waves per SIMD) of the FP32 inner-CG kernels
#ifdef TL_WPE32
#define TL_CGG_ATTR __attribute__((amdgpu_waves_per_eu(TL_WPE32)))
#else
#define TL_CGG_ATTR
#endif
// the same for the fused FP32 kernel
#ifdef TL_WPE_F32
#define TL_CGF_ATTR __attribute__((amdgpu_waves_per_eu(TL_WPE_F32)))
#else
#define TL_CGF_ATTR
#endif
// storage of the FP32 inner CG's cell coefficient gc: 16 bits (1) or FP32
// (0).  16-bit storage is IEEE half with a power-of-two scale when the
// coefficient's range fits half's normal range (ratio <= 2^28), else
// bfloat16.  Rounding a coefficient perturbs each face conductivity by a
// relative 2^-11 (half) or 2^-8 (bfloat16) at most, so the FP32 operator is
// spectrally equivalent to the FP64 one within that factor and the FP64
// residual refresh still converges to the FP64 operator's solution; the
// kernels read 2 bytes per cell instead of 4.
#ifndef TL_GC16
#define TL_GC16 1
#endif
#if TL_GC16
typedef unsigned short tl_gc_t;  // half or bfloat16 bits
#else
typedef float tl_gc_t;
#endif
enum GcFmt : int { GC_BF16 = 0, GC_F16 = 1 };
// blocks of the coefficient range reduction (k_gc_range)
#define TL_RANGE_BLOCKS 256
// non-temporal (streaming) vector loads / stores in the vectorised stencil
// kernels: each sweep reads and writes fields far larger than the caches.
// Measured at 8192^2 (A/B in one box, loads and stores together against
// neither): fused FP32 CG 4.68 / 4.78 -> 5.14 / 5.04 TB/s (odd / even t), FP32
// PW 4.65 -> 5.37, FP32 UR 5.03 / 5.09 -> 6.08 / 5.98, FP64 PW 5.23 -> 5.33,
// FP64 UR 5.17 -> 5.56, Chebyshev / PPCG / Jacobi steps +6-7%; loads alone
// made the fused kernel 4% slower.  The two-step solver kernels have their
// own switches: with both they ran 2-4% slower, with loads alone 6-8% slower,
// with stores alone within +-2%.
#ifndef TL_NT_LOAD
#define TL_NT_LOAD 1
#endif
#ifndef TL_NT_STORE
#define TL_NT_STORE 1
#endif
#ifndef TL_NT_2STEP_LD
#define TL_NT_2STEP_LD 0
#endif
#ifndef TL_NT_2STEP_ST
#define TL_NT_2STEP_ST 0
#endif

struct Geo {
    int nx, ny, hd;
    int X, Y;                  // nx + 2hd, ny + 2hd
    long pitch;                // elements per stored row
    long off;                  // element offset of interior (0, 0)
    int gx, gy, ty;            // stencil launch: gx*gy blocks of TB columns x ty rows
    int gx4, gy4, ty4;         // FP32 vector kernels: blocks of 4*TB columns x ty4 rows
    int gyp4, typ4;            // FP32 PW kernel: gx4*gyp4 blocks of 4*TB columns x typ4 rows
    int gx2, gy2, ty2;         // FP64 vector kernels: blocks of 2*TB columns x ty2 rows
    int gyf, tyf;              // fused FP32 kernel: gx4*gyf blocks of 4*TB columns x tyf rows
    int gyq, tyq;              // two-step PPCG kernel: gx2*gyq blocks of 2*TB columns x tyq rows
    int nbmax;                 // stride of the partial-sum arrays (max block count)
    size_t elems;              // elements per allocation
};

// Device-resident solver scalars.  Kernels read them at start and the last
// block of a reduction kernel writes them, so a timestep never round-trips
// through the host between iterations.
enum StopMode : int {
    MODE_CG_REF = 0,    // sqrt|rr| < eps                       (CG.jl:24)
    MODE_CG_REL = 1,    // sqrt|rr| <= tol_abs                  (relative rule)
    MODE_SWITCH = 2,    // CG phase of Cheby/PPCG: |rr| < eps, or the switch test of Cheby.jl:25-29
    MODE_IR = 3,        // FP32 CG: refresh trigger or candidate convergence
    MODE_PPCG = 4,      // PPCG outer iteration: |rr| < eps     (PPCG.jl:50)
    MODE_JACOBI = 5,    // |err| < eps                          (Jacobi.jl:26)
};
enum Why : int { WHY_NONE = 0, WHY_CONVERGED = 1, WHY_SWITCH = 2, WHY_MAXITERS = 3, WHY_REFRESH = 4,
                 WHY_BREAKDOWN = 5 };

struct Ctrl {
    double rro, pw, alpha, beta, err;
    double bb;          // sum u0^2
    double tol_abs;     // absolute stop threshold for MODE_CG_REL / MODE_IR
    double refresh_abs; // MODE_IR: refresh when sqrt(rr) drops below this
    double eps, epslim;
    double rel_tol;     // MODE_CG_REL / MODE_IR: tol_abs = rel_tol * sqrt(bb)
    double ir_delta;    // MODE_IR: refresh_abs = ir_delta * sqrt(rr_true)
    double sums[4];     // results of single-block finalizers (field summary, norms)
    double rrs[2];      // vectorised CG: rr(t) in rrs[t & 1], written once per iteration by
                        // block 0 of the kernel that sums it (or by the init/refresh finalizer)
    double alphas[2];   // alpha(t) in alphas[t & 1] (the FP32 inner CG applies two steps'
                        // x updates at once, at even t)
    int tg, tgn;        // hipGraph batches: first iteration of the running / next replay
    int iters;          // last completed iteration
    int maxiters;
    int mode;
    int presteps, errorswitch;
    int stop_iter;      // iterations t > stop_iter do not run (INT_MAX while running)
    int why;
    int status;
};

struct Unet;  // learned warm start (tl_unet.hip)

}  // namespace tl

struct tl_ctx {
    tl::Geo g;
    int device = 0;
    hipStream_t stream = nullptr;
    double xmin = 0, ymin = 0, dx = 1, dy = 1;
    // FP64 fields (reference Chunk fields; u and sd have ping-pong twins)
    double *f[TL_NFIELDS] = {};
    double *u_alt = nullptr, *sd_alt = nullptr, *guess = nullptr;
    double *r_alt1 = nullptr, *r_alt2 = nullptr;  // PPCG: r between two-step sweeps
    // FP32 fields of the mixed-precision solve
    // FP32 inner CG: cell coefficient gc = 1/w (tl_gc_t) and its face scale
    // (rx/2, ry/2), search direction, residual, correction
    tl::tl_gc_t *gcell = nullptr;
    int gc_fmt = 0;           // tl::GcFmt of gcell (16-bit storage)
    float gc_scale = 1.0f;    // gcell holds gc * gc_scale (a power of two)
    float *gc_range = nullptr;  // per-block min / max of gc (device, 2 * TL_RANGE_BLOCKS)
    float *p32 = nullptr, *r32 = nullptr, *x32 = nullptr;
    float hx32 = 0.0f, hy32 = 0.0f;
    // reductions
    double *partials = nullptr;
    size_t partials_len = 0;
    unsigned nb_w = 0, nb_r[2] = {0, 0}, nb_m = 0;  // block counts of the last producer of each partial array
    double *p_alt = nullptr;   // ping-pong twin of p for the fused CG kernels (FP64)
    double *cg_u = nullptr;    // the vector the fused FP64 UR adds alpha p to: u, or (relative stop
                               // rule) a correction added to u at every true-residual check
    float *p32b = nullptr;     // same for FP32
    // fused FP32 inner CG: second residual buffer (r ping-pongs, being read
    // at the neighbours), two sets of the four per-block sums, block counts
    float *r32b = nullptr;
    double *part_f = nullptr;
    unsigned nb_f[2] = {0, 0};
    tl::Ctrl *ctrl = nullptr;        // device
    tl::Ctrl *ctrl_h = nullptr;      // pinned host mirror
    double *hist_a = nullptr, *hist_b = nullptr;  // device CG alpha/beta history
    int hist_len = 0;
    int last_cg_iters = 0;
    // rasterisation
    tl_state *states_d = nullptr;
    std::string err;
    hipEvent_t ev[2] = {nullptr, nullptr};   // batch pipeline
    // per-launch event timing of selected kernels (tl_profile_begin)
    unsigned prof_mask = 0;
    unsigned prof_period = 1;   // bracket every prof_period-th launch of a kernel
    unsigned prof_seq[TL_K_COUNT] = {};  // launches per kernel since tl_profile_begin
    std::vector<hipEvent_t> prof_ev;          // pairs, reused
    std::vector<int> prof_tag;                // kernel id of each recorded pair
    size_t prof_used = 0;                     // events recorded since the last harvest
    std::vector<double> prof_ms[TL_K_COUNT];
    tl::Unet *unet = nullptr;
    bool recording = false;
    // learned warm start under TL_WARM_POLICY_BACKOFF: learned solves so far,
    // the one that attempts next, and the current gap after rejections
    long warm_solves = 0, warm_next = 0;
    int warm_gap = 1;
    // hipGraph batches of CG iterations (TL_LAUNCH_GRAPH), keyed by solve kind,
    // batch length and every launch argument that can change between solves
    struct GraphBatch {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<uintptr_t> key;
    };
    std::vector<GraphBatch> graphs;  // at most TL_GRAPH_CACHE entries
    unsigned graph_next = 0;         // entry replaced next when the cache is full
    bool capturing = false;
};

namespace tl {

// ------------------------------------------------------------ launchers
// All launchers enqueue on ctx->stream and return hipError_t.  Iteration
// kernels take the 1-based iteration number t; they do nothing once
// t > Ctrl::stop_iter.
hipError_t launch_init_ctrl(tl_ctx *c, int mode, int maxiters, double eps, double rel_tol, double ir_delta,
                            int presteps, int errorswitch, double epslim);
hipError_t launch_set_mode(tl_ctx *c, int mode, int maxiters);
hipError_t launch_set_state(tl_ctx *c, const tl_state *states_d, int nstates);
hipError_t launch_copy_full(tl_ctx *c, const double *a, double *b);
hipError_t launch_copy_vec(tl_ctx *c, const double *a, double *b);
hipError_t launch_copy_energy(tl_ctx *c);
// what k_init_coef writes besides kx, ky
enum { INIT_CG = 0, INIT_JACOBI = 1, INIT_COEF_ONLY = 2 };
hipError_t launch_init_coef(tl_ctx *c, double rx, double ry, int coefficient, int what);
hipError_t launch_halo(tl_ctx *c, unsigned mask, int depth, double *const *extra, int nextra);
hipError_t launch_cg_init_resid(tl_ctx *c, const double *guess);
hipError_t launch_resid_norm_slot(tl_ctx *c, const double *b, const double *x, int slot);
// tref > 0: the launch is being captured into a hipGraph batch whose first
// iteration is tref (t only selects the buffers; the kernel takes its
// iteration number from Ctrl::tg)
hipError_t launch_cg_pw64(tl_ctx *c, int t, int tref = 0);
hipError_t launch_cg_urv64(tl_ctx *c, int t, int tref = 0);
hipError_t launch_cg_pfinal64(tl_ctx *c, int t);
hipError_t launch_cg_w64(tl_ctx *c, int t);
hipError_t launch_cg_ur64(tl_ctx *c, int t);
hipError_t launch_cg_p64(tl_ctx *c, int t);
hipError_t launch_cg_w32(tl_ctx *c, int t, int tref = 0);
hipError_t launch_cg_ur32(tl_ctx *c, int t, int tref = 0);
hipError_t launch_graph_base(tl_ctx *c, int k);
hipError_t launch_graph_seed(tl_ctx *c, int t0);
hipError_t launch_residual(tl_ctx *c);
hipError_t launch_cg64_true_residual(tl_ctx *c, int T);
hipError_t launch_jacobi_residual(tl_ctx *c, const double *u, int t);
hipError_t launch_ir_residual(tl_ctx *c, int reset_p, int t_next, float *r_dst = nullptr);
hipError_t launch_cgf(tl_ctx *c, int t, int tref = 0);
hipError_t launch_cgf_dots(tl_ctx *c, int T);
float *cgf_r(tl_ctx *c, int t);
float *cgf_p(tl_ctx *c, int t);
hipError_t launch_cgp(tl_ctx *c, int t, bool first, int tref = 0);
float *cgp_r(tl_ctx *c, int T);
float *cgp_p(tl_ctx *c, int T);
hipError_t launch_apply(tl_ctx *c, const double *src, double *dst);
hipError_t launch_finalise(tl_ctx *c);
hipError_t launch_field_summary(tl_ctx *c);
hipError_t launch_sumsq_slot(tl_ctx *c, const double *v, int slot);
hipError_t launch_sumsq_rr(tl_ctx *c, const double *v, int t);
hipError_t launch_cheby_init(tl_ctx *c, double theta, const double *u_in, double *u_out);
hipError_t launch_cheby_iter(tl_ctx *c, double a, double b, const double *u_in, double *u_out, double *p, int norm,
                             int t);
hipError_t launch_cheby_iter2(tl_ctx *c, double a1, double b1, double a2, double b2, const double *u_in,
                              double *u_out, const double *p_in, double *p_out, int norm, int t);
hipError_t launch_ppcg_ur_sd(tl_ctx *c, double theta, int t);
hipError_t launch_ppcg_inner(tl_ctx *c, double a, double b, const double *sd_in, double *sd_out, int last, int t);
hipError_t launch_ppcg_inner2(tl_ctx *c, double a1, double b1, double a2, double b2, const double *sd_in,
                              double *sd_out, const double *r_in, double *r_out, int last, int t);
hipError_t launch_jacobi(tl_ctx *c, const double *u_in, double *u_out, int t, int defer);
hipError_t launch_ir_prepare(tl_ctx *c, double rx, double ry, int coefficient);
hipError_t launch_ir_accumulate(tl_ctx *c, int T, const float *pT = nullptr);
hipError_t launch_acc64(tl_ctx *c, double *u, double *d);

Geo make_geo(int nx, int ny, int hd);

}  // namespace tl
