// tl_unet.hip -- learned warm start: a BF16 U-Net on MFMA, trained on the
// device from the solver's own trajectories.  No counterpart in the
// reference; the network is defined by this file:
//
//   input: 16 channels (6 used), per cell of the grid, normalised
//          c0 = u_prev / s_u   c1 = log10(rho) / 2   c2 = energy / s_e
//          c3 = log10(1 + Kx) / 6   c4 = log10(1 + Ky) / 6
//          c5 = r0 / s_r with r0 = u_prev - A u_prev, the residual of the
//               previous-step guess (the step's right-hand side is u_prev)
//          s_u, s_e, s_r = max |u_prev|, |energy|, |r0| over the grid
//   enc1  conv3x3 16->16 ReLU, conv3x3 16->16 ReLU           (skip1, H x W)
//   avg-pool 2x2
//   enc2  conv3x3 16->32 ReLU, conv3x3 32->32 ReLU           (skip2, H/2)
//   avg-pool 2x2
//   mid   conv3x3 32->64 ReLU, conv3x3 64->64 ReLU           (H/4)
//   nearest x2, concat [up 64 | skip2 32] -> dec2 conv 96->32 ReLU, 32->32 ReLU
//   nearest x2, concat [up 32 | skip1 16] -> dec1 conv 48->16 ReLU, 16->16 ReLU
//   out   conv3x3 16->1 (linear): the predicted (u_new - u_prev) / s_r
// The change of u over a step is A^-1 r0, linear in r0: scaling by s_r makes
// every step of a trajectory look alike to the network.
//
// Activations are BF16 NHWC tensors with a one-cell zero border (the conv's
// zero padding needs no bounds checks); weights are BF16 copies of FP32
// master weights; every convolution accumulates in FP32 on the matrix cores
// (v_mfma_f32_16x16x16_bf16: a wave computes 16 pixels x 16 output channels
// per instruction, K = 16 input channels of one tap).  The input-gradient of
// a convolution is the same kernel on the output-gradient with flipped,
// transposed weights.  Weight gradients are FP32 sums over pixels (per-block
// partials reduced in fixed order), then Adam on the FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tl_internal.hpp"

namespace tl {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

// output rows per wave of the convolutions: the band of rows a wave walks
// (fewer on small tensors, so that a launch still has about TL_CONV_BLOCKS
// blocks); 16 and 64 measured within 2% of 32 at 8192^2
#ifndef TL_CONV_ROWS
#define TL_CONV_ROWS 32
#endif
#ifndef TL_CONV_BLOCKS
#define TL_CONV_BLOCKS 4096
#endif
// occupancy targets (waves per SIMD) of the convolutions with 16 and 32
// output channels per launch; the compiler otherwise spends registers down to
// 2 and 1 waves.  Measured at 8192^2: 16-channel launches at 2 waves (168
// VGPRs) make the inference 6% faster than at 3 (a few registers spilled)
#ifndef TL_CONV_WPE1
#define TL_CONV_WPE1 2
#endif
#ifndef TL_CONV_WPE2
#define TL_CONV_WPE2 2
#endif

constexpr int NL = 11;
struct LayerDef {
    int cin, cinp, cout, coutp, relu;
};
static const LayerDef LAY[NL] = {
    {6, 16, 16, 16, 1},  {16, 16, 16, 16, 1}, {16, 16, 32, 32, 1}, {32, 32, 32, 32, 1},
    {32, 32, 64, 64, 1}, {64, 64, 64, 64, 1}, {96, 96, 32, 32, 1}, {32, 32, 32, 32, 1},
    {48, 48, 16, 16, 1}, {16, 16, 16, 16, 1}, {16, 16, 1, 16, 0},
};

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}

// a bordered NHWC tensor view: element (b, y, x, c), -1 <= y <= H, -1 <= x <= W
struct TV {
    bf16 *p;
    int B, H, W, C, off;  // C = stored channels, off = first channel of the view
    __host__ __device__ long idx(int b, int y, int x) const
    {
        return (((long)b * (H + 2) + (y + 1)) * (W + 2) + (x + 1)) * C + off;
    }
};

// ------------------------------------------------------------- kernels

// Convolution 3x3, zero padding, as implicit GEMM on MFMA with the output
// channels on the M side.  A wave owns 64 pixels (4 groups of 16) of a column
// strip and walks a band of R output rows down the strip; a block is 4 waves
// side by side.  Each input row is loaded once per 16-channel chunk (its 3
// column shifts x 4 pixel groups: 12 fragments), the next (row, chunk)'s
// fragments requested before this one's products, and added into the three
// output rows it borders, kept in three rolling accumulator sets: input row r
// is row dy = 2 of output r-1's window (complete after it: stored), dy = 1 of
// output r and dy = 0 of output r+1.  Per chunk and tap: A = W[tap][co][ci]
// (16 co x 16 ci; lane (i, kq) holds ci 4kq..4kq+3 of co i), B =
// in[pixel][ci] (16 ci x 16 pixels; lane (j, kq) holds ci 4kq..4kq+3 of pixel
// j), D[co][pixel]: lane (j, kq) holds output channels 4kq..4kq+3 of pixel j,
// stored as one 8-byte vector.  A launch computes output channels co0 ..
// co0 + 16 NT - 1 of coutp on the pixel grid of `out`.
// Input channels [0, csplit) come from `in`, [csplit, cinp) from `skip`
// (channel c - csplit): a decoder's concatenation [up | skip] read in place.
// With upin, `in` has half the resolution and is read at (y/2, x/2), the
// nearest-neighbour upsample (its zero border lands on the output's border).
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT == 1 ? TL_CONV_WPE1 : TL_CONV_WPE2))) void k_conv(
    TV in, TV skip, int csplit, int upin, TV out, const bf16 *__restrict__ wgt, const float *__restrict__ bias,
    int cinp, int coutp, int co0, int relu, int R)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = lane & 15, kq = lane >> 4;
    const int y0 = blockIdx.y * R, b = blockIdx.z;
    const int x0 = (blockIdx.x * 4 + wave) * 64;
    // this launch's weights, [tap][16 NT output channels][cinp], copied into
    // the LDS once per block: every wave reads all of them for every row, and
    // from the caches they would be fetched by all CUs from the same lines
    extern __shared__ __attribute__((aligned(16))) bf16 wl[];
    {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const int per = cinp / 8, nvec = 9 * 16 * NT * per;  // 16-byte vectors
        for (int v = threadIdx.x; v < nvec; v += 256) {
            const int rowi = v / per, c8 = (v % per) * 8;
            const int tap = rowi / (16 * NT), co = rowi % (16 * NT);
            *(u4 *)(wl + (long)rowi * cinp + c8) = *(const u4 *)(wgt + ((long)tap * coutp + co0 + co) * cinp + c8);
        }
    }
    __syncthreads();
    if (x0 >= out.W) return;
    const int yout = min(y0 + R, out.H);  // output rows y0 .. yout-1
    int xa[4];  // this lane's pixel of each group (clamped: pixels past W are never stored)
#pragma unroll
    for (int m = 0; m < 4; ++m) xa[m] = min(x0 + 16 * m + j, out.W - 1);
    float bb[NT][4];  // requested first: their latency overlaps the main loop
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) bb[n][i] = bias ? bias[co0 + n * 16 + 4 * kq + i] : 0.f;
    f4 a0[4][NT], a1[4][NT], a2[4][NT];
    auto zero = [&](f4 (&a)[4][NT]) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) a[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    };
    zero(a0);
    zero(a1);
    zero(a2);
    auto store = [&](const f4 (&a)[4][NT], int y) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int co = co0 + n * 16 + 4 * kq;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int x = x0 + 16 * m + j;
                if (x < out.W) {
                    uint32_t w2[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        float v0 = a[m][n][2 * h] + bb[n][2 * h], v1 = a[m][n][2 * h + 1] + bb[n][2 * h + 1];
                        if (relu) {
                            v0 = fmaxf(v0, 0.f);
                            v1 = fmaxf(v1, 0.f);
                        }
                        w2[h] = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
                    }
                    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
                    *(u2 *)(out.p + out.idx(b, y, x) + co) = u2{w2[0], w2[1]};
                }
            }
        }
    };
    // the 12 fragments of input row r, chunk kc
    auto load = [&](s4 (&f)[3][4], int r, int kc) {
        if (kc < csplit) {
            const int ry = upin ? r >> 1 : r;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int x = xa[m] + dx - 1;
                    f[dx][m] = *(const s4 *)(in.p + in.idx(b, ry, upin ? x >> 1 : x) + kc + kq * 4);
                }
        } else {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    f[dx][m] = *(const s4 *)(skip.p + skip.idx(b, r, xa[m] + dx - 1) + kc - csplit + kq * 4);
        }
    };
    s4 fc[3][4], fn[3][4];
    load(fc, y0 - 1, 0);
    // input row r into the outputs r-1 (up), r (mid), r+1 (dn); then output
    // r-1 is complete
    auto row = [&](int r, f4 (&up)[4][NT], f4 (&mid)[4][NT], f4 (&dn)[4][NT]) {
        for (int kc = 0; kc < cinp; kc += 16) {
            // the next step: the next chunk of this row, or the next row's first
            const bool last = kc + 16 >= cinp;
            if (!last || r + 1 <= yout) load(fn, last ? r + 1 : r, last ? 0 : kc + 16);
            // keeps the following steps' loads below this step's products
            asm volatile("" ::: "memory");
            s4 wv[9][NT];
#pragma unroll
            for (int tap = 0; tap < 9; ++tap)
#pragma unroll
                for (int n = 0; n < NT; ++n) wv[tap][n] = *(const s4 *)(wl + (tap * 16 * NT + n * 16 + j) * cinp + kc + kq * 4);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int n = 0; n < NT; ++n)
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        dn[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wv[dx][n], fc[dx][m], dn[m][n], 0, 0, 0);
                        mid[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wv[3 + dx][n], fc[dx][m], mid[m][n], 0, 0, 0);
                        up[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wv[6 + dx][n], fc[dx][m], up[m][n], 0, 0, 0);
                    }
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int m = 0; m < 4; ++m) fc[dx][m] = fn[dx][m];
        }
        if (r - 1 >= y0 && r - 1 < yout) store(up, r - 1);
        zero(up);
    };
    // input rows y0-1 .. yout (row yout may be the zero bottom border); the
    // roles of the three accumulator sets rotate with the row
    for (int r = y0 - 1; r <= yout; r += 3) {
        row(r, a0, a1, a2);
        if (r + 1 > yout) break;
        row(r + 1, a1, a2, a0);
        if (r + 2 > yout) break;
        row(r + 2, a2, a0, a1);
    }
}

// 2x2 average pool: out (H/2) from in (H) view, 8 channels (16 bytes) per item
__global__ void k_pool(TV in, TV out, int C)
{
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const int C8 = C / 8;
    const long n = (long)out.B * out.H * out.W * C8;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C8) * 8;
        long r = t / C8;
        const int x = (int)(r % out.W);
        r /= out.W;
        const int y = (int)(r % out.H);
        const int b = (int)(r / out.H);
        const u4 q0 = *(const u4 *)(in.p + in.idx(b, 2 * y, 2 * x) + c);
        const u4 q1 = *(const u4 *)(in.p + in.idx(b, 2 * y, 2 * x + 1) + c);
        const u4 q2 = *(const u4 *)(in.p + in.idx(b, 2 * y + 1, 2 * x) + c);
        const u4 q3 = *(const u4 *)(in.p + in.idx(b, 2 * y + 1, 2 * x + 1) + c);
        u4 o;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t w = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int sh = 16 * h;
                const float s = bf2f((bf16)(q0[k] >> sh)) + bf2f((bf16)(q1[k] >> sh)) + bf2f((bf16)(q2[k] >> sh)) +
                                bf2f((bf16)(q3[k] >> sh));
                w |= (uint32_t)f2bf(0.25f * s) << sh;
            }
            o[k] = w;
        }
        *(u4 *)(out.p + out.idx(b, y, x) + c) = o;
    }
}

// nearest x2 upsample: out (2H) from in (H), 8 channels (16 bytes) per item
__global__ void k_up(TV in, TV out, int C)
{
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const int C8 = C / 8;
    const long n = (long)out.B * out.H * out.W * C8;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C8) * 8;
        long r = t / C8;
        const int x = (int)(r % out.W);
        r /= out.W;
        const int y = (int)(r % out.H);
        const int b = (int)(r / out.H);
        *(u4 *)(out.p + out.idx(b, y, x) + c) = *(const u4 *)(in.p + in.idx(b, y / 2, x / 2) + c);
    }
}

// backward of the upsample: out (H) = sum of the 2x2 block of in (2H)
__global__ void k_up_back(TV in, TV out, int C)
{
    const long n = (long)out.B * out.H * out.W * C;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t % C);
    long r = t / C;
    const int x = (int)(r % out.W);
    r /= out.W;
    const int y = (int)(r % out.H);
    const int b = (int)(r / out.H);
    float s = bf2f(in.p[in.idx(b, 2 * y, 2 * x) + c]) + bf2f(in.p[in.idx(b, 2 * y, 2 * x + 1) + c]) +
              bf2f(in.p[in.idx(b, 2 * y + 1, 2 * x) + c]) + bf2f(in.p[in.idx(b, 2 * y + 1, 2 * x + 1) + c]);
    out.p[out.idx(b, y, x) + c] = f2bf(s);
    }
}

// dz = (g [+ pool_back(gp)]) * (y > 0 if relu): the gradient at a conv's
// pre-activation; gp (H/2) is the gradient of the 2x2 pool that reads y
__global__ void k_dz(TV g, TV gp, int has_gp, TV y, int relu, TV dz, int C)
{
    const long n = (long)dz.B * dz.H * dz.W * C;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t % C);
    long r = t / C;
    const int x = (int)(r % dz.W);
    r /= dz.W;
    const int yy = (int)(r % dz.H);
    const int b = (int)(r / dz.H);
    float v = bf2f(g.p[g.idx(b, yy, x) + c]);
    if (has_gp) v += 0.25f * bf2f(gp.p[gp.idx(b, yy / 2, x / 2) + c]);
    if (relu && !(bf2f(y.p[y.idx(b, yy, x) + c]) > 0.f)) v = 0.f;
    dz.p[dz.idx(b, yy, x) + c] = f2bf(v);
    }
}

// weight-gradient partials: block (chunk) sums over its pixels, thread = one
// (co, ci) pair with its 9 taps: part[chunk][tap][co][ci]
__global__ __launch_bounds__(256) void k_dw(TV x, TV dz, int cinp, int coutp, int px_per_chunk,
                                            float *__restrict__ part)
{
    const int pair = blockIdx.x * 256 + threadIdx.x;
    const int chunk = blockIdx.y;
    const int npair = cinp * coutp;
    const long npx = (long)x.B * x.H * x.W;
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int co = pair / cinp, ci = pair % cinp;
    if (pair < npair) {
        const long p0 = (long)chunk * px_per_chunk;
        const long p1 = min(p0 + px_per_chunk, npx);
        for (long p = p0; p < p1; ++p) {
            const int xx = (int)(p % x.W);
            const long r = p / x.W;
            const int yy = (int)(r % x.H);
            const int b = (int)(r / x.H);
            const float g = bf2f(dz.p[dz.idx(b, yy, xx) + co]);
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int dy = tap / 3 - 1, dx = tap % 3 - 1;
                acc[tap] += g * bf2f(x.p[x.idx(b, yy + dy, xx + dx) + ci]);
            }
        }
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
            part[((long)chunk * 9 + tap) * npair + (long)co * cinp + ci] = acc[tap];
    }
}

// fixed-order sum of the partials into dw[tap][co][ci]
__global__ void k_dw_reduce(const float *__restrict__ part, int nchunk, long n, float *__restrict__ dw)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(long)c * n + i];
    dw[i] = s;
}

// bias gradient: db[co] = sum over pixels of dz[.., co] (one block per co)
__global__ __launch_bounds__(256) void k_db(TV dz, int co_real, float *__restrict__ db)
{
    const int co = blockIdx.x;
    const long npx = (long)dz.B * dz.H * dz.W;
    float s = 0.f;
    for (long p = threadIdx.x; p < npx; p += 256) {
        const int xx = (int)(p % dz.W);
        const long r = p / dz.W;
        const int yy = (int)(r % dz.H);
        const int b = (int)(r / dz.H);
        s += bf2f(dz.p[dz.idx(b, yy, xx) + co]);
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = co < co_real ? sh[0] : 0.f;
}

// FP32 master weights -> BF16 [tap][co][ci] and flipped/transposed
// BF16 [tap'][ci][co] = W[8 - tap'][co][ci]
__global__ void k_cvt_w(const float *__restrict__ w, int cinp, int coutp, bf16 *__restrict__ wb,
                        bf16 *__restrict__ wt)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = 9L * cinp * coutp;
    if (i >= n) return;
    const int ci = (int)(i % cinp);
    const int co = (int)((i / cinp) % coutp);
    const int tap = (int)(i / ((long)cinp * coutp));
    const bf16 h = f2bf(w[i]);
    wb[i] = h;
    wt[((long)(8 - tap) * cinp + ci) * coutp + co] = h;
}

// Adam (Kingma & Ba 2015) on FP32 parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// output gradient of the MSE loss on channel 0, masked to the scored
// region; partial sums of the squared error per block
__global__ __launch_bounds__(256) void k_loss(TV out, const float *__restrict__ tgt, const float *__restrict__ wmask,
                                              TV gout, float scale, float *__restrict__ part)
{
    const long n = (long)out.B * out.H * out.W;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.f;
    if (p < n) {
        const int xx = (int)(p % out.W);
        const long r = p / out.W;
        const int yy = (int)(r % out.H);
        const int b = (int)(r / out.H);
        const float wgt = wmask[p];
        const float d = (bf2f(out.p[out.idx(b, yy, xx)]) - tgt[p]) * wgt;
        e2 = d * d;
        const long gi = gout.idx(b, yy, xx);
        gout.p[gi] = f2bf(2.f * d * scale);
        for (int c = 1; c < gout.C; ++c) gout.p[gi + c] = 0;
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = e2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// face coefficients of a window: k[(y*W + x)*2 + {0,1}] = kx(x,y), ky(x,y)
__global__ void k_coefs(Geo g, const double *__restrict__ kx, const double *__restrict__ ky, float *__restrict__ k,
                        int W, int H, int x0, int y0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const long i = g.off + (long)(y0 + y) * g.pitch + (x0 + x);
    k[((long)y * W + x) * 2 + 0] = (float)kx[i];
    k[((long)y * W + x) * 2 + 1] = (float)ky[i];
}

// (A v)_i on a crop of the operator (smvp of kernels.jl:178-184): kx, ky
// per pixel (left / bottom faces); at a crop edge that is a domain edge the
// reflected neighbour equals v_i (edges: bit 0 left, 1 right, 2 bottom, 3 top);
// at other crop edges the missing neighbour is taken as 0
__device__ __forceinline__ float crop_apply(const float *__restrict__ v, const float *__restrict__ k, int x, int y,
                                            int W, int H, int edges)
{
    const long i = (long)y * W + x;
    const float vc = v[i];
    const float kxl = k[2 * i], kyd = k[2 * i + 1];
    const float kxr = x + 1 < W ? k[2 * (i + 1)] : kxl;
    const float kyu = y + 1 < H ? k[2 * (i + W) + 1] : kyd;
    const float vl = x > 0 ? v[i - 1] : ((edges & 1) ? vc : 0.f);
    const float vr = x + 1 < W ? v[i + 1] : ((edges & 2) ? vc : 0.f);
    const float vd = y > 0 ? v[i - W] : ((edges & 4) ? vc : 0.f);
    const float vu = y + 1 < H ? v[i + W] : ((edges & 8) ? vc : 0.f);
    return (1.f + kxl + kxr + kyd + kyu) * vc - kxl * vl - kxr * vr - kyd * vd - kyu * vu;
}

// residual loss, step 1: e = out - target on every pixel of the batch
__global__ void k_res_e(TV out, const float *__restrict__ tgt, float *__restrict__ e)
{
    const long n = (long)out.B * out.H * out.W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int xx = (int)(p % out.W);
    const long r = p / out.W;
    const int yy = (int)(r % out.H);
    const int b = (int)(r / out.H);
    e[p] = bf2f(out.p[out.idx(b, yy, xx)]) - tgt[p];
}

// step 2: r = mask * (A e) -- the residual of the predicted guess, in units
// of s_u, on the scored pixels; partial sums of r^2 per block
__global__ __launch_bounds__(256) void k_res_r(const float *__restrict__ e, const float *__restrict__ k,
                                               const float *__restrict__ mask, const int *__restrict__ edges, int B,
                                               int H, int W, float *__restrict__ r, float *__restrict__ part)
{
    const long n = (long)B * H * W;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    float r2 = 0.f;
    if (p < n) {
        const int x = (int)(p % W);
        const long q = p / W;
        const int y = (int)(q % H);
        const int b = (int)(q / H);
        const long o = (long)b * H * W;
        const float m = mask[p];
        const float v = m != 0.f ? m * crop_apply(e + o, k + 2 * o, x, y, W, H, edges[b]) : 0.f;
        r[p] = v;
        r2 = v * v;
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = r2;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// step 3: dL/d(out) = 2 scale A r (A symmetric), into channel 0 of gout
__global__ void k_res_g(const float *__restrict__ r, const float *__restrict__ k, const int *__restrict__ edges,
                        TV gout, float scale)
{
    const long n = (long)gout.B * gout.H * gout.W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int W = gout.W, H = gout.H;
    const int x = (int)(p % W);
    const long q = p / W;
    const int y = (int)(q % H);
    const int b = (int)(q / H);
    const long o = (long)b * H * W;
    const float g = 2.f * scale * crop_apply(r + o, k + 2 * o, x, y, W, H, edges[b]);
    const long gi = gout.idx(b, y, x);
    gout.p[gi] = f2bf(g);
    for (int c = 1; c < gout.C; ++c) gout.p[gi + c] = 0;
}

// energy-norm loss: e (already masked) ; r = A e ; partial sums of e . r
__global__ __launch_bounds__(256) void k_en_r(const float *__restrict__ e, const float *__restrict__ k,
                                              const int *__restrict__ edges, int B, int H, int W,
                                              float *__restrict__ r, float *__restrict__ part)
{
    const long n = (long)B * H * W;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    float s = 0.f;
    if (p < n) {
        const int x = (int)(p % W);
        const long q = p / W;
        const int y = (int)(q % H);
        const int b = (int)(q / H);
        const long o = (long)b * H * W;
        const float v = crop_apply(e + o, k + 2 * o, x, y, W, H, edges[b]);
        r[p] = v;
        s = e[p] * v;
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dL/d(out) = 2 scale mask (A (mask e)) into channel 0 of gout
__global__ void k_en_g(const float *__restrict__ r, const float *__restrict__ mask, TV gout, float scale)
{
    const long n = (long)gout.B * gout.H * gout.W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int W = gout.W, H = gout.H;
    const int x = (int)(p % W);
    const long q = p / W;
    const int y = (int)(q % H);
    const int b = (int)(q / H);
    const long gi = gout.idx(b, y, x);
    gout.p[gi] = f2bf(2.f * scale * mask[p] * r[p]);
    for (int c = 1; c < gout.C; ++c) gout.p[gi + c] = 0;
}

// masked error e = mask * (out - target)
__global__ void k_err_m(TV out, const float *__restrict__ tgt, const float *__restrict__ mask, float *__restrict__ e)
{
    const long n = (long)out.B * out.H * out.W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int xx = (int)(p % out.W);
    const long r = p / out.W;
    const int yy = (int)(r % out.H);
    const int b = (int)(r / out.H);
    e[p] = mask[p] * (bf2f(out.p[out.idx(b, yy, xx)]) - tgt[p]);
}

// per-grid max |u|, |energy|, |r0| (partials per block, then one block)
__global__ __launch_bounds__(256) void k_absmax3(Geo g, const double *__restrict__ u, const double *__restrict__ e,
                                                 const double *__restrict__ r0, float *__restrict__ part)
{
    float m[3] = {0.f, 0.f, 0.f};
    const long n = (long)g.nx * g.ny;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int x = (int)(t % g.nx), y = (int)(t / g.nx);
        const long i = g.off + (long)y * g.pitch + x;
        m[0] = fmaxf(m[0], (float)fabs(u[i]));
        m[1] = fmaxf(m[1], (float)fabs(e[i]));
        m[2] = fmaxf(m[2], (float)fabs(r0[i]));
    }
    __shared__ float sh[3][256];
    for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = m[a];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = fmaxf(sh[a][threadIdx.x], sh[a][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) part[3 * blockIdx.x + a] = sh[a][0];
}

__global__ void k_absmax_fin(const float *__restrict__ part, int nb, float *__restrict__ scale)
{
    if (threadIdx.x != 0) return;
    for (int a = 0; a < 3; ++a) {
        float v = 0.f;
        for (int i = 0; i < nb; ++i) v = fmaxf(v, part[3 * i + a]);
        scale[a] = v > 0.f ? v : 1.f;
    }
}

// r0 = b - A x with b = x = u_prev: the residual of the previous-step guess
__global__ void k_r0(Geo g, const double *__restrict__ u, const double *__restrict__ kx,
                     const double *__restrict__ ky, double *__restrict__ r0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.nx || y >= g.ny) return;
    const long i = g.off + (long)y * g.pitch + x;
    const double c = u[i];
    const double l = x > 0 ? u[i - 1] : c, r = x < g.nx - 1 ? u[i + 1] : c;
    const double d = y > 0 ? u[i - g.pitch] : c, up = y < g.ny - 1 ? u[i + g.pitch] : c;
    const double kxl = kx[i], kxr = kx[i + 1], kyd = ky[i], kyu = ky[i + g.pitch];
    const double au = (1.0 + kxr + kxl + kyu + kyd) * c - (kxr * r + kxl * l) - (kyu * up + kyd * d);
    r0[i] = c - au;
}

// network input for the window [x0, x0+W) x [y0, y0+H) of the grid (batch b)
__global__ void k_input(Geo g, const double *__restrict__ u, const double *__restrict__ rho,
                        const double *__restrict__ e, const double *__restrict__ kx, const double *__restrict__ ky,
                        const double *__restrict__ r0, const float *__restrict__ scale, TV in, int b, int x0, int y0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= in.W || y >= in.H) return;
    const long i = g.off + (long)(y0 + y) * g.pitch + (x0 + x);
    bf16 *o = in.p + in.idx(b, y, x);
    o[0] = f2bf((float)(u[i] / scale[0]));
    o[1] = f2bf((float)(log10(rho[i]) * 0.5));
    o[2] = f2bf((float)(e[i] / scale[1]));
    o[3] = f2bf((float)(log10(1.0 + kx[i]) / 6.0));
    o[4] = f2bf((float)(log10(1.0 + ky[i]) / 6.0));
    o[5] = f2bf((float)(r0[i] / scale[2]));
    for (int c = 6; c < 16; ++c) o[c] = 0;
}

// training target (u_new - u_prev) / s_r for the same window
__global__ void k_target(Geo g, const double *__restrict__ unew, const double *__restrict__ uprev,
                         const float *__restrict__ scale, float *__restrict__ tgt, int W, int H, int x0, int y0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const long i = g.off + (long)(y0 + y) * g.pitch + (x0 + x);
    tgt[(long)y * W + x] = (float)((unew[i] - uprev[i]) / scale[2]);
}

// guess = u_prev + s_r * out (FP64) over the interior
__global__ void k_guess(Geo g, const double *__restrict__ uprev, TV out, const float *__restrict__ scale,
                        double *__restrict__ guess)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.nx || y >= g.ny) return;
    const long i = g.off + (long)y * g.pitch + x;
    guess[i] = uprev[i] + (double)scale[2] * (double)bf2f(out.p[out.idx(0, y, x)]);
}

// ------------------------------------------------------------- host side

struct Tensor {
    bf16 *p = nullptr;
    int B = 0, H = 0, W = 0, C = 0;
    size_t bytes() const { return (size_t)B * (H + 2) * (W + 2) * C * sizeof(bf16); }
    TV v(int off = 0) const { return TV{p, B, H, W, C, off}; }
};

struct Acts {
    int B = 0, H = 0, W = 0;
    bool grads = false;
    // forward
    Tensor x0, e1a, c1, p1, e2a, c2, p2, e3a, e3b, d2a, d2b, d1a, d1b, out;
    Tensor e1b, e2b;  // the skip tensors on their own (forward without concatenation)
    // backward
    Tensor g_out, g_d1b, g_d1a, g_c1, g_d2b, g_d2a, g_c2, g_e3b, g_e3a, g_p2, g_e2a, g_p1, g_e1a, g_e1b, g_e2b, dz;
};

}  // namespace tl

struct tl::Unet {
    float *w[NL] = {}, *b[NL] = {};
    float *mw[NL] = {}, *vw[NL] = {}, *mb[NL] = {}, *vb[NL] = {};
    float *gw[NL] = {}, *gb[NL] = {};
    bf16 *wb[NL] = {}, *wt[NL] = {};
    bool weights_set = false;
    int adam_t = 0;
    Acts acts;
    float *scale = nullptr;      // [3]: s_u, s_e, s_r
    float *mx_part = nullptr;    // absmax partials
    float *dw_part = nullptr;
    size_t dw_part_len = 0;
    float *loss_part = nullptr;
    size_t loss_part_len = 0;
    // dataset of training samples: inputs (bordered, 16 ch), target, loss mask
    bf16 *ds_x = nullptr;
    float *ds_y = nullptr, *ds_m = nullptr, *ds_k = nullptr;
    std::vector<int> ds_edges;
    int ds_n = 0, ds_cap = 0, ds_crop = 0, ds_per_step = 0;
    std::mt19937_64 rng{12345};
    // batch buffers
    float *bt_y = nullptr, *bt_m = nullptr, *bt_k = nullptr, *bt_e = nullptr, *bt_r = nullptr;
    int *bt_edges = nullptr;
    int bt_cap = 0;
};

namespace tl {

static size_t wcount(int l) { return 9UL * LAY[l].cinp * LAY[l].coutp; }

#define UCHK(c, expr)                                                      \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess) {                                            \
            (c)->err = std::string("unet: ") + #expr + ": " + hipGetErrorString(e_); \
            return TL_E_HIP;                                               \
        }                                                                  \
    } while (0)

static int ualloc(tl_ctx *c, void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        c->err = "unet: device allocation failed";
        *p = nullptr;
        return TL_E_NOMEM;
    }
    if (hipMemsetAsync(*p, 0, bytes, c->stream) != hipSuccess) return TL_E_HIP;
    return TL_OK;
}

void unet_destroy(tl_ctx *c);

static int unet_alloc(tl_ctx *c, Unet *u)
{
    for (int l = 0; l < NL; ++l) {
        const size_t nw = wcount(l), nb = LAY[l].coutp;
        float **fw[] = {&u->w[l], &u->mw[l], &u->vw[l], &u->gw[l]};
        for (float **q : fw)
            if (int rc = ualloc(c, (void **)q, nw * sizeof(float))) return rc;
        float **fb[] = {&u->b[l], &u->mb[l], &u->vb[l], &u->gb[l]};
        for (float **q : fb)
            if (int rc = ualloc(c, (void **)q, nb * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->wb[l], nw * sizeof(bf16))) return rc;
        if (int rc = ualloc(c, (void **)&u->wt[l], nw * sizeof(bf16))) return rc;
    }
    if (int rc = ualloc(c, (void **)&u->scale, 4 * sizeof(float))) return rc;
    return ualloc(c, (void **)&u->mx_part, 3 * 1024 * sizeof(float));
}

// the context's network, created on first use; a failed creation leaves none
static int unet_get(tl_ctx *c, Unet **out)
{
    if (!c->unet) {
        c->unet = new Unet();
        if (int rc = unet_alloc(c, c->unet)) {
            unet_destroy(c);
            return rc;
        }
    }
    *out = c->unet;
    return TL_OK;
}

static void tfree(Tensor &t)
{
    if (t.p) (void)hipFree(t.p);
    t = Tensor();
}

static int talloc(tl_ctx *c, Tensor &t, int B, int H, int W, int C)
{
    t.B = B;
    t.H = H;
    t.W = W;
    t.C = C;
    return ualloc(c, (void **)&t.p, t.bytes());
}

static void acts_free(Acts &a)
{
    Tensor *all[] = {&a.x0,    &a.e1a,   &a.c1,    &a.p1,    &a.e2a,   &a.c2,    &a.p2,    &a.e3a,
                     &a.e3b,   &a.d2a,   &a.d2b,   &a.d1a,   &a.d1b,   &a.out,   &a.g_out, &a.g_d1b,
                     &a.g_d1a, &a.g_c1,  &a.g_d2b, &a.g_d2a, &a.g_c2,  &a.g_e3b, &a.g_e3a, &a.g_p2,
                     &a.g_e2a, &a.g_p1,  &a.g_e1a, &a.g_e1b, &a.g_e2b, &a.dz,    &a.e1b,   &a.e2b};
    for (Tensor *t : all) tfree(*t);
    a.B = a.H = a.W = 0;
}

// activations (and gradients) for B x H x W; H, W multiples of 4
static int acts_ensure(tl_ctx *c, Unet *u, int B, int H, int W, bool grads)
{
    Acts &a = u->acts;
    if (a.B == B && a.H == H && a.W == W && (a.grads || !grads)) return TL_OK;
    acts_free(a);
    a.B = B;
    a.H = H;
    a.W = W;
    a.grads = grads;
    const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    struct {
        Tensor *t;
        int h, w, ch;
    } f[] = {{&a.x0, H, W, 16},   {&a.e1a, H, W, 16},    {&a.c1, H, W, 48},    {&a.p1, H2, W2, 16},
             {&a.e2a, H2, W2, 32}, {&a.c2, H2, W2, 96},  {&a.p2, H4, W4, 32},  {&a.e3a, H4, W4, 64},
             {&a.e3b, H4, W4, 64}, {&a.d2a, H2, W2, 32}, {&a.d2b, H2, W2, 32}, {&a.d1a, H, W, 16},
             {&a.d1b, H, W, 16},   {&a.out, H, W, 16},   {&a.e1b, H, W, 16},   {&a.e2b, H2, W2, 32}};
    for (auto &q : f)
        if (int rc = talloc(c, *q.t, B, q.h, q.w, q.ch)) {
            acts_free(a);
            return rc;
        }
    if (grads) {
        struct {
            Tensor *t;
            int h, w, ch;
        } g[] = {{&a.g_out, H, W, 16},   {&a.g_d1b, H, W, 16},   {&a.g_d1a, H, W, 16},  {&a.g_c1, H, W, 48},
                 {&a.g_d2b, H2, W2, 32}, {&a.g_d2a, H2, W2, 32}, {&a.g_c2, H2, W2, 96}, {&a.g_e3b, H4, W4, 64},
                 {&a.g_e3a, H4, W4, 64}, {&a.g_p2, H4, W4, 32},  {&a.g_e2a, H2, W2, 32}, {&a.g_p1, H2, W2, 16},
                 {&a.g_e1a, H, W, 16},   {&a.g_e1b, H, W, 16},   {&a.g_e2b, H2, W2, 32}, {&a.dz, H, W, 64}};
        for (auto &q : g)
            if (int rc = talloc(c, *q.t, B, q.h, q.w, q.ch)) {
                acts_free(a);
                return rc;
            }
    }
    return TL_OK;
}

static unsigned blocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
// grid of a grid-stride kernel: a dispatch counts its work-items in 32 bits,
// so full-grid tensors of 16384^2 x 64 channels need the loop
static unsigned gs_blocks(long n, int bs = 256) { return (unsigned)std::min<long>((n + bs - 1) / bs, 1L << 20); }

// out = conv(in) (same size); or, with skip, out = conv([up2(in) | skip])
// with csplit channels from in (upsampled 2x when upin)
static hipError_t conv(tl_ctx *c, const TV &in, const TV &out, const bf16 *w, const float *bias, int cinp, int coutp,
                       int relu, const TV *skip = nullptr, int csplit = 0, int upin = 0)
{
    if (coutp % 16 || coutp < 16 || cinp % 16 || cinp < 16 || cinp > 96) return hipErrorInvalidValue;
    if (skip && (csplit % 16 || csplit <= 0 || csplit >= cinp)) return hipErrorInvalidValue;
    // 32 output channels per launch (16 for the last 16 of 48): the three
    // rolling accumulator sets stay at 96 registers
    const int gx = (out.W + 255) / 256;
    int R = TL_CONV_ROWS;
    while (R > 4 && (long)gx * ((out.H + R - 1) / R) * out.B < TL_CONV_BLOCKS) R /= 2;
    const dim3 grid(gx, (out.H + R - 1) / R, out.B);
    const TV &sk = skip ? *skip : in;
    const int cs = skip ? csplit : cinp;
    for (int co0 = 0; co0 < coutp; co0 += 32) {
        const int nt = coutp - co0 >= 32 ? 2 : 1;
        const size_t lds = (size_t)9 * 16 * nt * cinp * sizeof(bf16);  // at most 55 KB (96 -> 32)
        if (nt == 2)
            k_conv<2><<<grid, 256, lds, c->stream>>>(in, sk, cs, upin, out, w, bias, cinp, coutp, co0, relu, R);
        else
            k_conv<1><<<grid, 256, lds, c->stream>>>(in, sk, cs, upin, out, w, bias, cinp, coutp, co0, relu, R);
    }
    return hipGetLastError();
}

static hipError_t fwd_layer(tl_ctx *c, Unet *u, int l, const TV &in, const TV &out)
{
    return conv(c, in, out, u->wb[l], u->b[l], LAY[l].cinp, LAY[l].coutp, LAY[l].relu);
}

static int refresh_weights(tl_ctx *c, Unet *u)
{
    for (int l = 0; l < NL; ++l) {
        const long n = (long)wcount(l);
        k_cvt_w<<<blocks(n), 256, 0, c->stream>>>(u->w[l], LAY[l].cinp, LAY[l].coutp, u->wb[l], u->wt[l]);
    }
    UCHK(c, hipGetLastError());
    return TL_OK;
}

// forward pass over acts.x0
// The forward pass.  With concat (a backward pass follows), the decoders'
// inputs are materialised as the concatenations c1 = [up(d2b) | e1b] and
// c2 = [up(e3b) | e2b] that the weight gradients read; without, the decoder
// convolutions read the low-resolution tensor (upsampled on the fly) and the
// skip tensor in place, and the pools read the skip tensors whole.
static int forward(tl_ctx *c, Unet *u, bool concat)
{
    Acts &a = u->acts;
    const long n1 = (long)a.B * a.H * a.W, n2 = n1 / 4;
    UCHK(c, fwd_layer(c, u, 0, a.x0.v(), a.e1a.v()));
    if (concat) {
        UCHK(c, fwd_layer(c, u, 1, a.e1a.v(), a.c1.v(32)));
        k_pool<<<gs_blocks(n2 * 2), 256, 0, c->stream>>>(a.c1.v(32), a.p1.v(), 16);
        UCHK(c, fwd_layer(c, u, 2, a.p1.v(), a.e2a.v()));
        UCHK(c, fwd_layer(c, u, 3, a.e2a.v(), a.c2.v(64)));
        k_pool<<<gs_blocks(n2 / 4 * 4), 256, 0, c->stream>>>(a.c2.v(64), a.p2.v(), 32);
    } else {
        UCHK(c, fwd_layer(c, u, 1, a.e1a.v(), a.e1b.v()));
        k_pool<<<gs_blocks(n2 * 2), 256, 0, c->stream>>>(a.e1b.v(), a.p1.v(), 16);
        UCHK(c, fwd_layer(c, u, 2, a.p1.v(), a.e2a.v()));
        UCHK(c, fwd_layer(c, u, 3, a.e2a.v(), a.e2b.v()));
        k_pool<<<gs_blocks(n2 / 4 * 4), 256, 0, c->stream>>>(a.e2b.v(), a.p2.v(), 32);
    }
    UCHK(c, fwd_layer(c, u, 4, a.p2.v(), a.e3a.v()));
    UCHK(c, fwd_layer(c, u, 5, a.e3a.v(), a.e3b.v()));
    if (concat) {
        k_up<<<gs_blocks(n2 * 8), 256, 0, c->stream>>>(a.e3b.v(), a.c2.v(0), 64);
        UCHK(c, fwd_layer(c, u, 6, a.c2.v(), a.d2a.v()));
    } else {
        const TV sk = a.e2b.v();
        UCHK(c, conv(c, a.e3b.v(), a.d2a.v(), u->wb[6], u->b[6], LAY[6].cinp, LAY[6].coutp, LAY[6].relu, &sk, 64, 1));
    }
    UCHK(c, fwd_layer(c, u, 7, a.d2a.v(), a.d2b.v()));
    if (concat) {
        k_up<<<gs_blocks(n1 * 4), 256, 0, c->stream>>>(a.d2b.v(), a.c1.v(0), 32);
        UCHK(c, fwd_layer(c, u, 8, a.c1.v(), a.d1a.v()));
    } else {
        const TV sk = a.e1b.v();
        UCHK(c, conv(c, a.d2b.v(), a.d1a.v(), u->wb[8], u->b[8], LAY[8].cinp, LAY[8].coutp, LAY[8].relu, &sk, 32, 1));
    }
    UCHK(c, fwd_layer(c, u, 9, a.d1a.v(), a.d1b.v()));
    UCHK(c, fwd_layer(c, u, 10, a.d1b.v(), a.out.v()));
    return TL_OK;
}

// backward of layer l: g = dL/d(output), y = output, x = input; writes the
// weight/bias gradients and (unless gx is null) dL/d(input) into gx
static int back_layer(tl_ctx *c, Unet *u, int l, const TV &g, const TV &gp, int has_gp, const TV &y, const TV &x,
                      const TV *gx)
{
    Acts &a = u->acts;
    const LayerDef &L = LAY[l];
    TV dz{a.dz.p, y.B, y.H, y.W, L.coutp, 0};
    const long npx = (long)y.B * y.H * y.W;
    // the buffer is shared by all layers: clear it so its border is zero
    UCHK(c, hipMemsetAsync(a.dz.p, 0, (size_t)y.B * (y.H + 2) * (y.W + 2) * L.coutp * sizeof(bf16), c->stream));
    k_dz<<<gs_blocks(npx * L.coutp), 256, 0, c->stream>>>(g, gp, has_gp, y, L.relu, dz, L.coutp);
    // weight gradient: partials over pixel chunks, then a fixed-order sum
    const int px_chunk = 1024;
    const int nchunk = (int)((npx + px_chunk - 1) / px_chunk);
    const long npair = (long)L.cinp * L.coutp;
    const size_t need = (size_t)nchunk * 9 * npair;
    if (need > u->dw_part_len) {
        if (u->dw_part) (void)hipFree(u->dw_part);
        u->dw_part = nullptr;
        u->dw_part_len = 0;
        if (int rc = ualloc(c, (void **)&u->dw_part, need * sizeof(float))) return rc;
        u->dw_part_len = need;
    }
    k_dw<<<dim3(blocks(npair), nchunk), 256, 0, c->stream>>>(x, dz, L.cinp, L.coutp, px_chunk, u->dw_part);
    k_dw_reduce<<<blocks(9 * npair), 256, 0, c->stream>>>(u->dw_part, nchunk, 9 * npair, u->gw[l]);
    k_db<<<L.coutp, 256, 0, c->stream>>>(dz, L.cout, u->gb[l]);
    if (gx) UCHK(c, conv(c, dz, *gx, u->wt[l], nullptr, L.coutp, L.cinp, 0));
    UCHK(c, hipGetLastError());
    return TL_OK;
}

static int backward(tl_ctx *c, Unet *u)
{
    Acts &a = u->acts;
    const long n1 = (long)a.B * a.H * a.W, n2 = n1 / 4;
    const TV none{};
    TV t;
#define RB(expr)                   \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != TL_OK) return rc_; \
    } while (0)
    t = a.g_d1b.v();
    RB(back_layer(c, u, 10, a.g_out.v(), none, 0, a.out.v(), a.d1b.v(), &t));
    t = a.g_d1a.v();
    RB(back_layer(c, u, 9, a.g_d1b.v(), none, 0, a.d1b.v(), a.d1a.v(), &t));
    t = a.g_c1.v();
    RB(back_layer(c, u, 8, a.g_d1a.v(), none, 0, a.d1a.v(), a.c1.v(), &t));
    k_up_back<<<gs_blocks(n2 * 32), 256, 0, c->stream>>>(a.g_c1.v(0), a.g_d2b.v(), 32);
    t = a.g_d2a.v();
    RB(back_layer(c, u, 7, a.g_d2b.v(), none, 0, a.d2b.v(), a.d2a.v(), &t));
    t = a.g_c2.v();
    RB(back_layer(c, u, 6, a.g_d2a.v(), none, 0, a.d2a.v(), a.c2.v(), &t));
    k_up_back<<<gs_blocks(n2 / 4 * 64), 256, 0, c->stream>>>(a.g_c2.v(0), a.g_e3b.v(), 64);
    t = a.g_e3a.v();
    RB(back_layer(c, u, 5, a.g_e3b.v(), none, 0, a.e3b.v(), a.e3a.v(), &t));
    t = a.g_p2.v();
    RB(back_layer(c, u, 4, a.g_e3a.v(), none, 0, a.e3a.v(), a.p2.v(), &t));
    // skip2: gradient from the concat plus the pool's
    t = a.g_e2a.v();
    RB(back_layer(c, u, 3, a.g_c2.v(64), a.g_p2.v(), 1, a.c2.v(64), a.e2a.v(), &t));
    t = a.g_p1.v();
    RB(back_layer(c, u, 2, a.g_e2a.v(), none, 0, a.e2a.v(), a.p1.v(), &t));
    t = a.g_e1a.v();
    RB(back_layer(c, u, 1, a.g_c1.v(32), a.g_p1.v(), 1, a.c1.v(32), a.e1a.v(), &t));
    RB(back_layer(c, u, 0, a.g_e1a.v(), none, 0, a.e1a.v(), a.x0.v(), nullptr));
#undef RB
    return TL_OK;
}

static int adam_step(tl_ctx *c, Unet *u, float lr)
{
    u->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
    for (int l = 0; l < NL; ++l) {
        const long nw = (long)wcount(l), nb = LAY[l].coutp;
        k_adam<<<blocks(nw), 256, 0, c->stream>>>(u->w[l], u->gw[l], u->mw[l], u->vw[l], nw, lr, b1, b2, eps, c1, c2);
        k_adam<<<blocks(nb), 256, 0, c->stream>>>(u->b[l], u->gb[l], u->mb[l], u->vb[l], nb, lr, b1, b2, eps, c1, c2);
    }
    UCHK(c, hipGetLastError());
    return TL_OK;
}

// host-side He initialisation (deterministic), padded layout
static int init_weights(tl_ctx *c, Unet *u, uint64_t seed)
{
    std::mt19937_64 g(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int l = 0; l < NL; ++l) {
        const LayerDef &L = LAY[l];
        std::vector<float> w(wcount(l), 0.f);
        const float sd = std::sqrt((L.relu ? 2.f : 1.f) / (9.f * L.cin));
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci) w[((size_t)t * L.coutp + co) * L.cinp + ci] = nd(g) * sd;
        UCHK(c, hipMemcpyAsync(u->w[l], w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        UCHK(c, hipMemsetAsync(u->b[l], 0, L.coutp * sizeof(float), c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));
    }
    u->weights_set = true;
    return refresh_weights(c, u);
}

// r0 of u_prev into `r0`, then the grid scales s_u, s_e, s_r
static int grid_scales(tl_ctx *c, Unet *u, const double *uprev, const double *energy, double *r0)
{
    const int nb = 1024;
    k_r0<<<dim3(blocks(c->g.nx), c->g.ny), 256, 0, c->stream>>>(c->g, uprev, c->f[TL_F_KX], c->f[TL_F_KY], r0);
    k_absmax3<<<nb, 256, 0, c->stream>>>(c->g, uprev, energy, r0, u->mx_part);
    k_absmax_fin<<<1, 64, 0, c->stream>>>(u->mx_part, nb, u->scale);
    UCHK(c, hipGetLastError());
    return TL_OK;
}

// ------------------------------------------------------------ entry points

int unet_predict_guess(tl_ctx *c, double, double, int, double *guess_dev)
{
    Unet *u = c->unet;
    if (!u || !u->weights_set) {
        c->err = "learned warm start: no network weights loaded (tl_unet_set_weights)";
        return TL_E_STATE;
    }
    const Geo &g = c->g;
    const int H = (g.ny + 3) / 4 * 4, W = (g.nx + 3) / 4 * 4;
    if (H != g.ny || W != g.nx) {
        c->err = "learned warm start needs nx, ny multiples of 4";
        return TL_E_INVALID;
    }
    if (int rc = acts_ensure(c, u, 1, H, W, false)) return rc;
    // r0 goes to the u_alt scratch field (free during a CG solve)
    if (int rc = grid_scales(c, u, c->f[TL_F_U], c->f[TL_F_ENERGY], c->u_alt)) return rc;
    k_input<<<dim3(blocks(W), H), 256, 0, c->stream>>>(g, c->f[TL_F_U], c->f[TL_F_DENSITY], c->f[TL_F_ENERGY],
                                                        c->f[TL_F_KX], c->f[TL_F_KY], c->u_alt, u->scale,
                                                        u->acts.x0.v(), 0, 0, 0);
    UCHK(c, hipGetLastError());
    if (int rc = forward(c, u, false)) return rc;
    k_guess<<<dim3(blocks(g.nx), g.ny), 256, 0, c->stream>>>(g, c->f[TL_F_U], u->acts.out.v(), u->scale, guess_dev);
    UCHK(c, hipGetLastError());
    return TL_OK;
}

void unet_destroy(tl_ctx *c)
{
    Unet *u = c->unet;
    if (!u) return;
    acts_free(u->acts);
    for (int l = 0; l < NL; ++l) {
        void *ps[] = {u->w[l], u->b[l], u->mw[l], u->vw[l], u->mb[l], u->vb[l], u->gw[l], u->gb[l], u->wb[l], u->wt[l]};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
    void *ps[] = {u->scale, u->mx_part, u->dw_part, u->loss_part, u->ds_x, u->ds_y, u->ds_m, u->ds_k,
                  u->bt_y, u->bt_m, u->bt_k, u->bt_e, u->bt_r, u->bt_edges};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete u;
    c->unet = nullptr;
}

// After a solve in recording mode: crops of (inputs of this step, target).
// u0 holds u_prev, u the solution, energy the previous energy.
int unet_record_step(tl_ctx *c, double, double, int)
{
    Unet *u;
    if (int rc = unet_get(c, &u)) return rc;
    const Geo &g = c->g;
    const int crop = u->ds_crop;
    if (crop <= 0 || crop > g.nx || crop > g.ny || u->ds_n >= u->ds_cap) return TL_OK;
    if (int rc = grid_scales(c, u, c->f[TL_F_U0], c->f[TL_F_ENERGY], c->u_alt)) return rc;
    const int per_step = std::min(u->ds_cap - u->ds_n, u->ds_per_step);
    const int margin = 16;
    std::vector<float> mask((size_t)crop * crop);
    for (int k = 0; k < per_step; ++k) {
        const int x0 = (int)(u->rng() % (uint64_t)(g.nx - crop + 1));
        const int y0 = (int)(u->rng() % (uint64_t)(g.ny - crop + 1));
        const int s = u->ds_n++;
        TV in{u->ds_x, u->ds_cap, crop, crop, 16, 0};
        k_input<<<dim3(blocks(crop), crop), 256, 0, c->stream>>>(g, c->f[TL_F_U0], c->f[TL_F_DENSITY],
                                                                 c->f[TL_F_ENERGY], c->f[TL_F_KX], c->f[TL_F_KY],
                                                                 c->u_alt, u->scale, in, s, x0, y0);
        k_target<<<dim3(blocks(crop), crop), 256, 0, c->stream>>>(g, c->f[TL_F_U], c->f[TL_F_U0], u->scale,
                                                                  u->ds_y + (size_t)s * crop * crop, crop, crop, x0,
                                                                  y0);
        k_coefs<<<dim3(blocks(crop), crop), 256, 0, c->stream>>>(g, c->f[TL_F_KX], c->f[TL_F_KY],
                                                                 u->ds_k + (size_t)s * 2 * crop * crop, crop, crop,
                                                                 x0, y0);
        u->ds_edges[s] = (x0 == 0 ? 1 : 0) | (x0 + crop == g.nx ? 2 : 0) | (y0 == 0 ? 4 : 0) |
                         (y0 + crop == g.ny ? 8 : 0);
        // score pixels away from crop edges that are not domain edges (there
        // the zero padding of the crop differs from the full-grid input)
        for (int y = 0; y < crop; ++y)
            for (int x = 0; x < crop; ++x) {
                const bool okx = (x >= margin || x0 == 0) && (x < crop - margin || x0 + crop == g.nx);
                const bool oky = (y >= margin || y0 == 0) && (y < crop - margin || y0 + crop == g.ny);
                mask[(size_t)y * crop + x] = (okx && oky) ? 1.f : 0.f;
            }
        UCHK(c, hipMemcpyAsync(u->ds_m + (size_t)s * crop * crop, mask.data(), mask.size() * sizeof(float),
                               hipMemcpyHostToDevice, c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));  // mask buffer reuse
    }
    UCHK(c, hipGetLastError());
    return TL_OK;
}

}  // namespace tl

using namespace tl;

extern "C" int tl_unet_param_count(void)
{
    int n = 0;
    for (int l = 0; l < NL; ++l) n += 9 * LAY[l].cin * LAY[l].cout + LAY[l].cout;
    return n;
}

// flat real layout per layer: W[tap][co][ci] (cout x cin real), then b[co]
extern "C" int tl_unet_set_weights(tl_ctx *c, const float *flat)
{
    if (!c) return TL_E_INVALID;
    UCHK(c, hipSetDevice(c->device));
    Unet *u;
    if (int rc = unet_get(c, &u)) return rc;
    if (!flat) return init_weights(c, u, 0x7ea1ea5ULL);
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        const LayerDef &L = LAY[l];
        std::vector<float> w(wcount(l), 0.f), b(L.coutp, 0.f);
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci) w[((size_t)t * L.coutp + co) * L.cinp + ci] = flat[o++];
        for (int co = 0; co < L.cout; ++co) b[co] = flat[o++];
        UCHK(c, hipMemcpyAsync(u->w[l], w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        UCHK(c, hipMemcpyAsync(u->b[l], b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));
    }
    u->weights_set = true;
    u->adam_t = 0;
    return refresh_weights(c, u);
}

extern "C" int tl_unet_get_weights(tl_ctx *c, float *flat)
{
    if (!c || !flat || !c->unet) return TL_E_INVALID;
    Unet *u = c->unet;
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        const LayerDef &L = LAY[l];
        std::vector<float> w(wcount(l)), b(L.coutp);
        UCHK(c, hipMemcpyAsync(w.data(), u->w[l], w.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        UCHK(c, hipMemcpyAsync(b.data(), u->b[l], b.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci) flat[o++] = w[((size_t)t * L.coutp + co) * L.cinp + ci];
        for (int co = 0; co < L.cout; ++co) flat[o++] = b[co];
    }
    return TL_OK;
}

// capacity > 0: record crops_per_step crops of 128 x 128 cells after every
// solve, up to `capacity` samples in all; capacity 0 stops recording (the
// dataset is kept)
extern "C" int tl_unet_record(tl_ctx *c, int capacity, int crops_per_step)
{
    if (!c || capacity < 0 || crops_per_step < 0) return TL_E_INVALID;
    UCHK(c, hipSetDevice(c->device));
    Unet *u;
    if (int rc = unet_get(c, &u)) return rc;
    u->ds_per_step = crops_per_step;
    if (capacity > 0 && !u->ds_x) {
        u->ds_crop = 128;
        u->ds_cap = capacity;
        const size_t px = (size_t)u->ds_crop * u->ds_crop;
        const size_t xb = (size_t)u->ds_cap * (u->ds_crop + 2) * (u->ds_crop + 2) * 16 * sizeof(bf16);
        int rc = ualloc(c, (void **)&u->ds_x, xb);
        if (rc == TL_OK) rc = ualloc(c, (void **)&u->ds_y, (size_t)u->ds_cap * px * sizeof(float));
        if (rc == TL_OK) rc = ualloc(c, (void **)&u->ds_m, (size_t)u->ds_cap * px * sizeof(float));
        if (rc == TL_OK) rc = ualloc(c, (void **)&u->ds_k, (size_t)u->ds_cap * 2 * px * sizeof(float));
        if (rc != TL_OK) {  // no partial dataset: the next call allocates it again
            void **ds[] = {(void **)&u->ds_x, (void **)&u->ds_y, (void **)&u->ds_m, (void **)&u->ds_k};
            for (void **q : ds) {
                if (*q) (void)hipFree(*q);
                *q = nullptr;
            }
            u->ds_cap = 0;
            return rc;
        }
        u->ds_edges.assign(u->ds_cap, 0);
    }
    c->recording = capacity > 0;
    return TL_OK;
}

extern "C" int tl_unet_dataset_size(tl_ctx *c) { return (c && c->unet) ? c->unet->ds_n : 0; }

// Adam on minibatches of recorded crops; loss_out[i] = mean squared error of
// step i over the scored pixels
extern "C" int tl_unet_train(tl_ctx *c, const tl_train_params *tp, float *loss_out)
{
    if (!c || !tp || tp->steps < 1 || tp->batch < 1) return TL_E_INVALID;
    UCHK(c, hipSetDevice(c->device));
    Unet *u;
    if (int rc = unet_get(c, &u)) return rc;
    if (u->ds_n < 1) {
        c->err = "tl_unet_train: empty dataset (tl_unet_record, then solve)";
        return TL_E_STATE;
    }
    if (!u->weights_set)
        if (int rc = init_weights(c, u, 0x7ea1ea5ULL + (uint64_t)tp->seed)) return rc;
    const int crop = u->ds_crop, B = tp->batch;
    const size_t px = (size_t)crop * crop;
    if (int rc = acts_ensure(c, u, B, crop, crop, true)) return rc;
    if (u->bt_cap < B) {
        void **old[] = {(void **)&u->bt_y, (void **)&u->bt_m, (void **)&u->bt_k,
                        (void **)&u->bt_e, (void **)&u->bt_r, (void **)&u->bt_edges};
        for (void **q : old) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        u->bt_cap = 0;
        if (int rc = ualloc(c, (void **)&u->bt_y, B * px * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->bt_m, B * px * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->bt_k, 2 * B * px * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->bt_e, B * px * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->bt_r, B * px * sizeof(float))) return rc;
        if (int rc = ualloc(c, (void **)&u->bt_edges, B * sizeof(int))) return rc;
        u->bt_cap = B;
    }
    const int loss_kind = tp->reserved[0];  // 0 residual, 1 mse, 2 energy norm
    std::vector<int> edges(B);
    const long npx = (long)B * px;
    const unsigned nlb = blocks(npx);
    if ((size_t)nlb > u->loss_part_len) {
        if (u->loss_part) (void)hipFree(u->loss_part);
        u->loss_part = nullptr;
        u->loss_part_len = 0;
        if (int rc = ualloc(c, (void **)&u->loss_part, nlb * sizeof(float))) return rc;
        u->loss_part_len = nlb;
    }
    std::mt19937_64 g(0x5eed0000ULL + (uint64_t)tp->seed);
    const size_t xs = (size_t)(crop + 2) * (crop + 2) * 16;
    std::vector<float> lp(nlb), mk(npx);
    for (int step = 0; step < tp->steps; ++step) {
        // gather a batch
        for (int b = 0; b < B; ++b) {
            const int s = (int)(g() % (uint64_t)u->ds_n);
            edges[b] = u->ds_edges[s];
            UCHK(c, hipMemcpyAsync(u->bt_k + 2 * b * px, u->ds_k + (size_t)s * 2 * px, 2 * px * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
            UCHK(c, hipMemcpyAsync(u->acts.x0.p + b * xs, u->ds_x + (size_t)s * xs, xs * sizeof(bf16),
                                   hipMemcpyDeviceToDevice, c->stream));
            UCHK(c, hipMemcpyAsync(u->bt_y + b * px, u->ds_y + (size_t)s * px, px * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
            UCHK(c, hipMemcpyAsync(u->bt_m + b * px, u->ds_m + (size_t)s * px, px * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
        }
        if (int rc = forward(c, u, true)) return rc;
        // scale = 1 / (number of scored pixels) (from the masks of the batch)
        UCHK(c, hipMemcpyAsync(mk.data(), u->bt_m, npx * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));
        double nsc = 0;
        for (long i = 0; i < npx; ++i) nsc += mk[i];
        const float scale = nsc > 0 ? (float)(1.0 / nsc) : 0.f;
        if (loss_kind == 2) {
            // energy norm of the error: L = scale * e^T A e with e = mask (out - target)
            UCHK(c, hipMemcpyAsync(u->bt_edges, edges.data(), B * sizeof(int), hipMemcpyHostToDevice, c->stream));
            k_err_m<<<nlb, 256, 0, c->stream>>>(u->acts.out.v(), u->bt_y, u->bt_m, u->bt_e);
            k_en_r<<<nlb, 256, 0, c->stream>>>(u->bt_e, u->bt_k, u->bt_edges, B, crop, crop, u->bt_r, u->loss_part);
            k_en_g<<<nlb, 256, 0, c->stream>>>(u->bt_r, u->bt_m, u->acts.g_out.v(), scale);
        } else if (loss_kind == 0) {
            // minimise the residual of the guess: L = scale * sum_scored (A (out - target))^2
            UCHK(c, hipMemcpyAsync(u->bt_edges, edges.data(), B * sizeof(int), hipMemcpyHostToDevice, c->stream));
            k_res_e<<<nlb, 256, 0, c->stream>>>(u->acts.out.v(), u->bt_y, u->bt_e);
            k_res_r<<<nlb, 256, 0, c->stream>>>(u->bt_e, u->bt_k, u->bt_m, u->bt_edges, B, crop, crop, u->bt_r,
                                                 u->loss_part);
            k_res_g<<<nlb, 256, 0, c->stream>>>(u->bt_r, u->bt_k, u->bt_edges, u->acts.g_out.v(), scale);
        } else {
            k_loss<<<nlb, 256, 0, c->stream>>>(u->acts.out.v(), u->bt_y, u->bt_m, u->acts.g_out.v(), scale,
                                               u->loss_part);
        }
        UCHK(c, hipGetLastError());
        if (int rc = backward(c, u)) return rc;
        if (int rc = adam_step(c, u, tp->lr)) return rc;
        if (int rc = refresh_weights(c, u)) return rc;
        if (loss_out) {
            UCHK(c, hipMemcpyAsync(lp.data(), u->loss_part, nlb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            UCHK(c, hipStreamSynchronize(c->stream));
            double s = 0;
            for (float v : lp) s += v;
            loss_out[step] = (float)(s * scale);
        }
    }
    UCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

// prediction of the next u (FP64) into the TL_F_MI field, from the current
// u (energy*density after CG.init!) -- as the warm start computes it
extern "C" int tl_unet_predict(tl_ctx *c, double rx, double ry, int coefficient)
{
    if (!c) return TL_E_INVALID;
    UCHK(c, hipSetDevice(c->device));
    if (int rc = launch_init_coef(c, rx, ry, coefficient, INIT_CG)) return rc == hipSuccess ? TL_OK : TL_E_HIP;
    if (int rc = unet_predict_guess(c, rx, ry, coefficient, c->f[TL_F_MI])) return rc;
    UCHK(c, hipStreamSynchronize(c->stream));
    return TL_OK;
}

// Diagnostics for the parity tests: forward / loss+gradient of the network
// on host inputs x[B][H][W][6] (FP32, already normalised), target[B][H][W].
extern "C" int tl_unet_eval(tl_ctx *c, const float *x, const float *target, int B, int H, int W, float *out,
                            float *loss, float *grads)
{
    if (!c || !x || B < 1 || H < 4 || W < 4 || H % 4 || W % 4) return TL_E_INVALID;
    UCHK(c, hipSetDevice(c->device));
    Unet *u;
    if (int rc = unet_get(c, &u)) return rc;
    if (!u->weights_set) {
        c->err = "tl_unet_eval: no weights";
        return TL_E_STATE;
    }
    const bool want_grad = target && grads;
    if (int rc = acts_ensure(c, u, B, H, W, want_grad)) return rc;
    const size_t npx = (size_t)B * H * W;
    std::vector<bf16> hx(u->acts.x0.bytes() / sizeof(bf16), 0);
    auto tobf = [](float f) {
        uint32_t v;
        std::memcpy(&v, &f, 4);
        v += 0x7FFFu + ((v >> 16) & 1u);
        return (bf16)(v >> 16);
    };
    for (int b = 0; b < B; ++b)
        for (int y = 0; y < H; ++y)
            for (int xx = 0; xx < W; ++xx)
                for (int ch = 0; ch < 6; ++ch)
                    hx[((((size_t)b * (H + 2) + y + 1) * (W + 2)) + xx + 1) * 16 + ch] =
                        tobf(x[(((size_t)b * H + y) * W + xx) * 6 + ch]);
    UCHK(c, hipMemcpyAsync(u->acts.x0.p, hx.data(), hx.size() * sizeof(bf16), hipMemcpyHostToDevice, c->stream));
    if (int rc = forward(c, u, want_grad)) return rc;
    std::vector<bf16> ho(u->acts.out.bytes() / sizeof(bf16));
    UCHK(c, hipMemcpyAsync(ho.data(), u->acts.out.p, ho.size() * sizeof(bf16), hipMemcpyDeviceToHost, c->stream));
    UCHK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b)
        for (int y = 0; y < H; ++y)
            for (int xx = 0; xx < W; ++xx) {
                const uint32_t v = (uint32_t)ho[((((size_t)b * (H + 2) + y + 1) * (W + 2)) + xx + 1) * 16] << 16;
                float f;
                std::memcpy(&f, &v, 4);
                if (out) out[((size_t)b * H + y) * W + xx] = f;
            }
    if (!want_grad) return TL_OK;
    // scratch for the loss, freed on every return path
    struct Scratch {
        float *dy = nullptr, *dm = nullptr, *lp = nullptr;
        ~Scratch()
        {
            for (float *q : {dy, dm, lp})
                if (q) (void)hipFree(q);
        }
    } sc;
    const unsigned nlb = blocks((long)npx);
    if (hipMalloc((void **)&sc.dy, npx * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&sc.dm, npx * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&sc.lp, nlb * sizeof(float)) != hipSuccess) {
        c->err = "unet: device allocation failed";
        return TL_E_NOMEM;
    }
    float *dy = sc.dy, *dm = sc.dm, *lp = sc.lp;
    std::vector<float> ones(npx, 1.f), lph(nlb);
    UCHK(c, hipMemcpyAsync(dy, target, npx * sizeof(float), hipMemcpyHostToDevice, c->stream));
    UCHK(c, hipMemcpyAsync(dm, ones.data(), npx * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const float scale = 1.f / (float)npx;
    k_loss<<<nlb, 256, 0, c->stream>>>(u->acts.out.v(), dy, dm, u->acts.g_out.v(), scale, lp);
    UCHK(c, hipGetLastError());
    if (int rc = backward(c, u)) return rc;
    UCHK(c, hipMemcpyAsync(lph.data(), lp, nlb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    UCHK(c, hipStreamSynchronize(c->stream));
    double s = 0;
    for (float v : lph) s += v;
    if (loss) *loss = (float)(s * scale);
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        const LayerDef &L = LAY[l];
        std::vector<float> w(wcount(l)), b(L.coutp);
        UCHK(c, hipMemcpyAsync(w.data(), u->gw[l], w.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        UCHK(c, hipMemcpyAsync(b.data(), u->gb[l], b.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        UCHK(c, hipStreamSynchronize(c->stream));
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci) grads[o++] = w[((size_t)t * L.coutp + co) * L.cinp + ci];
        for (int co = 0; co < L.cout; ++co) grads[o++] = b[co];
    }
    return TL_OK;
}
