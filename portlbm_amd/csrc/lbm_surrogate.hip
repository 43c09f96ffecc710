// Warm-start surrogate: forward, backward and Adam kernels of a 14-conv
// encoder-decoder, and the executor that trains it on a device-resident
// dataset (include/lbm_surrogate.h).  FP32 throughout, FP64 only for the
// loss sums.  The network has no counterpart in the reference.
//
// Kernels (NCHW float32, 256-thread blocks):
//   k_conv3x3_f32  implicit GEMM on the matrix cores in FP32
//                  (v_mfma_f32_16x16x4f32): 4 output rows x 16 pixels per
//                  block (one row per wave) for all output channels; the input
//                  tile with its 1-node halo staged in LDS; weights packed
//                  [Cout16][K16], K = 9 Cin in the weights' own (c, tap) order.
//                  The data gradient is the same kernel on flipped, transposed
//                  weights.
//   k_conv_wgrad   16 x 16 (out, in) channel pairs per block, each thread one
//                  pair's 9 taps + bias over a sliding 3x3 register window;
//                  per-slice partials summed in two fixed-order stages
//                  (32 slice groups in parallel, then the groups).
//   pooling, upsampling, ReLU mask, loss and Adam: one thread per element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_error.hpp"
#include "lbm_hip.h"
#include "lbm_surrogate.h"

namespace
{

constexpr int kConvs = LBM_SURROGATE_CONVS;
constexpr int kCin[kConvs] = {4, 16, 16, 32, 32, 48, 48, 64, 64, 64, 48, 32, 16, 16};
constexpr int kCout[kConvs] = {16, 16, 32, 32, 48, 48, 64, 64, 64, 48, 32, 16, 16, 3};

constexpr int kTile = 16;                    // output tile edge (pixels)
constexpr int kHalo = kTile + 2;             // input tile edge with halo
constexpr int kHaloStride = kHalo * kHalo + 1;  // LDS floats per channel (odd: spreads banks)
constexpr int kWgT = 16;                     // channel tile of the weight gradient
constexpr int kGStride = kTile * kTile + 1;  // LDS floats per gradient channel

constexpr double kPi = 3.14159265358979323846;

inline int round16(int c) { return (c + 15) / 16 * 16; }

// Resolution level (0 = full, 4 = 1/16) at which conv k runs.
inline int conv_level(int k) { return k < 8 ? k / 2 : k == 8 ? 4 : k == 13 ? 0 : 12 - k; }

// ---- FP32 convolution on the matrix cores ---------------------------------
// y[b][o][p] = act(bias[o] + sum_k Wm[o][k] X[k][p]), k = c*9 + tap, with
// v_mfma_f32_16x16x4f32 (A 16x4, B 4x16, FP32 products and accumulation).
// Lane layout (the same as the BF16 instruction's, one element per lane):
// A[i][kk] in lane i + 16 kk, B[kk][j] in lane j + 16 kk, D[i][j] in lane
// j + 16 (i/4), element i%4.  Which k an operand slot kk carries is free as
// long as A and B agree: in a 16-wide chunk at kc, MFMA s of the chunk gives
// slot kk the k = kc + 4 kk + s, so that each lane reads its four A values
// as one 16-byte load.  A 256-thread block computes 16 pixels of 4 rows (one
// row per wave) for all output channels (at most 64: four 16-row tiles).
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaRows = 4;  // output rows per block (one per wave)
constexpr int kMfmaCols = 16;
constexpr int kMaxCoTiles = 4;  // Cout <= 64

__global__ __launch_bounds__(256) void k_conv3x3_f32(const float *__restrict__ x, const float *__restrict__ wm,
                                                     const float *__restrict__ bias, float *__restrict__ y, int Cin,
                                                     int Cout, int Kp, int H, int W, int tiles_x, int relu)
{
    constexpr int TR = kMfmaRows + 2, TC = kMfmaCols + 2;
    extern __shared__ float tile_f[];  // Cin x TR x TC
    const int b = blockIdx.y;
    const int x0 = int(blockIdx.x % unsigned(tiles_x)) * kMfmaCols;
    const int y0 = int(blockIdx.x / unsigned(tiles_x)) * kMfmaRows;
    const float *xb = x + size_t(b) * Cin * H * W;
    for (int i = threadIdx.x; i < Cin * TR * TC; i += 256)
    {
        const int c = i / (TR * TC), r = (i / TC) % TR, col = i % TC;
        const int yy = y0 - 1 + r, xx = x0 - 1 + col;
        float v = 0.f;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = xb[(size_t(c) * H + yy) * W + xx];
        tile_f[i] = v;
    }
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int j = lane % 16, kq = 4 * (lane / 16);
    const int co_tiles = (Cout + 15) / 16;
    const int K = 9 * Cin;
    f32x4 acc[kMaxCoTiles];
#pragma unroll
    for (int t = 0; t < kMaxCoTiles; ++t)
        acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Kp; kc += 16)
    {
        float bv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s)
        {
            const int k = kc + kq + s;
            float v = 0.f;
            if (k < K)
            {
                const int c = k / 9, tap = k - 9 * c;
                v = tile_f[(c * TR + wave + tap / 3) * TC + j + tap % 3];
            }
            bv[s] = v;
        }
#pragma unroll
        for (int t = 0; t < kMaxCoTiles; ++t)
        {
            if (t < co_tiles)
            {
                const f32x4 av = *reinterpret_cast<const f32x4 *>(wm + size_t(t * 16 + j) * Kp + kc + kq);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s], acc[t], 0, 0, 0);
            }
        }
    }
    const int oy = y0 + wave, ox = x0 + j;
    if (oy >= H || ox >= W)
        return;
#pragma unroll
    for (int t = 0; t < kMaxCoTiles; ++t)
    {
        if (t >= co_tiles)
            break;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int co = t * 16 + kq + v;
            if (co < Cout)
            {
                float r = acc[t][v] + (bias ? bias[co] : 0.f);
                if (relu)
                    r = r > 0.f ? r : 0.f;
                y[((size_t(b) * Cout + co) * H + oy) * W + ox] = r;
            }
        }
    }
}

// Partial weight/bias gradient of one slice of (sample, tile) pairs:
// partial[((slice*CoutP + o)*CinP + c)*10 + k] = sum g[b,o,i,j] x[b,c,i+ky-1,j+kx-1] (k < 9),
// slot 9 = sum g[b,o,i,j].
__global__ __launch_bounds__(256) void k_conv_wgrad(const float *__restrict__ x, const float *__restrict__ g,
                                                    float *__restrict__ partial, int B, int Cin, int Cout, int CinP,
                                                    int CoutP, int H, int W, int tiles_x, int tiles_per_image,
                                                    int slices)
{
    __shared__ float gs[kWgT * kGStride];
    __shared__ float xs[kWgT * kHaloStride];
    const int ci_tiles = CinP / kWgT;
    const int co0 = int(blockIdx.x / unsigned(ci_tiles)) * kWgT, ci0 = int(blockIdx.x % unsigned(ci_tiles)) * kWgT;
    const int ol = threadIdx.x / kWgT, cl = threadIdx.x % kWgT;
    float acc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k)
        acc[k] = 0.f;
    const int total = B * tiles_per_image;
    for (int t = blockIdx.y; t < total; t += slices)
    {
        const int b = t / tiles_per_image, tt = t % tiles_per_image;
        const int tx0 = (tt % tiles_x) * kTile, ty0 = (tt / tiles_x) * kTile;
        __syncthreads();
        for (int i = threadIdx.x; i < kWgT * kTile * kTile; i += 256)
        {
            const int c = i / (kTile * kTile), p = i % (kTile * kTile);
            const int yy = ty0 + p / kTile, xx = tx0 + p % kTile;
            float v = 0.f;
            if (co0 + c < Cout && yy < H && xx < W)
                v = g[((size_t(b) * Cout + co0 + c) * H + yy) * W + xx];
            gs[c * kGStride + p] = v;
        }
        for (int i = threadIdx.x; i < kWgT * kHalo * kHalo; i += 256)
        {
            const int c = i / (kHalo * kHalo), r = i % (kHalo * kHalo);
            const int yy = ty0 + r / kHalo - 1, xx = tx0 + r % kHalo - 1;
            float v = 0.f;
            if (ci0 + c < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = x[((size_t(b) * Cin + ci0 + c) * H + yy) * W + xx];
            xs[c * kHaloStride + r] = v;
        }
        __syncthreads();
        const float *xr = xs + cl * kHaloStride;
        const float *gr = gs + ol * kGStride;
        for (int r = 0; r < kTile; ++r)
        {
            float w0[3], w1[3], w2[3];
            w0[0] = xr[r * kHalo];
            w0[1] = xr[r * kHalo + 1];
            w1[0] = xr[(r + 1) * kHalo];
            w1[1] = xr[(r + 1) * kHalo + 1];
            w2[0] = xr[(r + 2) * kHalo];
            w2[1] = xr[(r + 2) * kHalo + 1];
#pragma unroll 4
            for (int c = 0; c < kTile; ++c)
            {
                w0[2] = xr[r * kHalo + c + 2];
                w1[2] = xr[(r + 1) * kHalo + c + 2];
                w2[2] = xr[(r + 2) * kHalo + c + 2];
                const float gv = gr[r * kTile + c];
                acc[0] = fmaf(gv, w0[0], acc[0]);
                acc[1] = fmaf(gv, w0[1], acc[1]);
                acc[2] = fmaf(gv, w0[2], acc[2]);
                acc[3] = fmaf(gv, w1[0], acc[3]);
                acc[4] = fmaf(gv, w1[1], acc[4]);
                acc[5] = fmaf(gv, w1[2], acc[5]);
                acc[6] = fmaf(gv, w2[0], acc[6]);
                acc[7] = fmaf(gv, w2[1], acc[7]);
                acc[8] = fmaf(gv, w2[2], acc[8]);
                acc[9] += gv;
                w0[0] = w0[1];
                w0[1] = w0[2];
                w1[0] = w1[1];
                w1[1] = w1[2];
                w2[0] = w2[1];
                w2[1] = w2[2];
            }
        }
    }
    float *out = partial + ((size_t(blockIdx.y) * CoutP + co0 + ol) * CinP + ci0 + cl) * 10;
#pragma unroll
    for (int k = 0; k < 10; ++k)
        out[k] = acc[k];
}

// Weight-gradient sums over the slices in two stages, each in a fixed order
// (bitwise reproducible): stage 1 sums group g's slices [g*per, (g+1)*per)
// of one weight per thread into part2[g][weight] (FP64); stage 2 adds the
// groups.  Weight index w < Cout*Cin*9 is (o, c, k), then Cout biases (slot 9
// of channel 0).
constexpr int kReduceGroups = 32;

__device__ __forceinline__ size_t wgrad_slot(int w, int Cin, int Cout, int CinP)
{
    const int nw = Cout * Cin * 9;
    int o, c, k;
    if (w < nw)
    {
        o = w / (Cin * 9);
        c = (w / 9) % Cin;
        k = w % 9;
    }
    else
    {
        o = w - nw;
        c = 0;
        k = 9;
    }
    return (size_t(o) * CinP + c) * 10 + k;
}

__global__ void k_wgrad_reduce(const float *__restrict__ partial, double *__restrict__ part2, int Cin, int Cout,
                               int CinP, int CoutP, int slices, int per)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int nall = Cout * Cin * 9 + Cout;
    if (w >= nall)
        return;
    const size_t slot = wgrad_slot(w, Cin, Cout, CinP);
    const size_t stride = size_t(CoutP) * CinP * 10;
    const int s0 = blockIdx.y * per, s1 = min(slices, s0 + per);
    double s = 0.0;
    for (int sl = s0; sl < s1; ++sl)
        s += double(partial[size_t(sl) * stride + slot]);
    part2[size_t(blockIdx.y) * nall + w] = s;
}

__global__ void k_wgrad_final(const double *__restrict__ part2, float *__restrict__ gw, float *__restrict__ gb,
                              int Cin, int Cout, int groups)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = Cout * Cin * 9, nall = nw + Cout;
    if (w >= nall)
        return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g)
        s += part2[size_t(g) * nall + w];
    if (w < nw)
        gw[w] = float(s);
    else
        gb[w - nw] = float(s);
}

// ---- BF16 inference: implicit-GEMM convolution on the matrix cores --------
// out[o][p] = sum_k Wb[o][k] X[k][p], k = c*9 + tap (the FP32 weight order),
// with v_mfma_f32_16x16x16bf16_1k (FP32 accumulation).  Lane layout of that
// instruction, checked on the device: A[i][k] in lane i + 16 (k/4), element
// k%4; B[k][j] in lane j + 16 (k/4), element k%4; D[i][j] in lane
// j + 16 (i/4), element i%4.  A 256-thread block computes 16 pixels of 4 rows
// (one row per wave) for all output channels: the input tile with its halo
// is staged in LDS as BF16 (round to nearest even), each wave builds the
// im2col fragment of a 16-wide K chunk once and reuses it for every 16-channel
// output tile; weights come packed as BF16 [Cout16][K16] (zero padded).
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return short((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

__global__ __launch_bounds__(256) void k_conv3x3_bf16(const float *__restrict__ x, const short *__restrict__ wb,
                                                      const float *__restrict__ bias, float *__restrict__ y, int Cin,
                                                      int Cout, int Kp, int H, int W, int tiles_x, int relu)
{
    constexpr int TR = kMfmaRows + 2, TC = kMfmaCols + 2;
    extern __shared__ short tile_bf[];  // Cin x TR x TC
    const int b = blockIdx.y;
    const int x0 = int(blockIdx.x % unsigned(tiles_x)) * kMfmaCols;
    const int y0 = int(blockIdx.x / unsigned(tiles_x)) * kMfmaRows;
    const float *xb = x + size_t(b) * Cin * H * W;
    for (int i = threadIdx.x; i < Cin * TR * TC; i += 256)
    {
        const int c = i / (TR * TC), r = (i / TC) % TR, col = i % TC;
        const int yy = y0 - 1 + r, xx = x0 - 1 + col;
        float v = 0.f;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = xb[(size_t(c) * H + yy) * W + xx];
        tile_bf[i] = f32_to_bf16_rne(v);
    }
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int j = lane % 16, kq = 4 * (lane / 16);
    const int co_tiles = (Cout + 15) / 16;
    const int K = 9 * Cin;
    f32x4 acc[kMaxCoTiles];
#pragma unroll
    for (int t = 0; t < kMaxCoTiles; ++t)
        acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Kp; kc += 16)
    {
        bf16x4 bv;
#pragma unroll
        for (int it = 0; it < 4; ++it)
        {
            const int k = kc + kq + it;
            short v = 0;
            if (k < K)
            {
                const int c = k / 9, tap = k - 9 * (k / 9);
                v = tile_bf[(c * TR + wave + tap / 3) * TC + j + tap % 3];
            }
            bv[it] = v;
        }
#pragma unroll
        for (int t = 0; t < kMaxCoTiles; ++t)
        {
            if (t < co_tiles)
            {
                const bf16x4 av = *reinterpret_cast<const bf16x4 *>(wb + size_t(t * 16 + j) * Kp + kc + kq);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
    const int oy = y0 + wave, ox = x0 + j;
    if (oy >= H || ox >= W)
        return;
#pragma unroll
    for (int t = 0; t < kMaxCoTiles; ++t)
    {
        if (t >= co_tiles)
            break;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int co = t * 16 + kq + v;
            if (co < Cout)
            {
                float r = acc[t][v] + bias[co];
                if (relu)
                    r = r > 0.f ? r : 0.f;
                y[((size_t(b) * Cout + co) * H + oy) * W + ox] = r;
            }
        }
    }
}

// wb[o*Kp + k] = bf16(w[o*9Cin + k]) for k < 9 Cin (padding stays zero)
__global__ void k_pack_bf16(const float *__restrict__ w, short *__restrict__ wb, int Cin, int Cout, int Kp)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int K = 9 * Cin;
    if (idx >= Cout * K)
        return;
    const int o = idx / K, k = idx % K;
    wb[size_t(o) * Kp + k] = f32_to_bf16_rne(w[idx]);
}

// wm[o*Kp + c*9 + k] = w[o][c][k] (Kp = round16(9 Cin));
// wt[c*KpT + o*9 + 8-k] = w[o][c][k] (KpT = round16(9 Cout)): the data
// gradient's weights, input and output channels swapped, taps flipped.
// Padding rows and columns stay zero.
__global__ void k_pack(const float *__restrict__ w, float *__restrict__ wm, float *__restrict__ wt, int Cin, int Cout,
                       int Kp, int KpT)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Cout * Cin * 9)
        return;
    const int o = idx / (Cin * 9), c = (idx / 9) % Cin, k = idx % 9;
    wm[size_t(o) * Kp + c * 9 + k] = w[idx];
    wt[size_t(c) * KpT + o * 9 + (8 - k)] = w[idx];
}

__global__ void k_pool2(const float *__restrict__ x, float *__restrict__ y, int BC, int H, int W)
{
    const int h = H / 2, w = W / 2;
    const size_t n = size_t(BC) * h * w;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t bc = i / (size_t(h) * w);
        const int r = int(i % (size_t(h) * w)), yi = r / w, xi = r % w;
        const float *p = x + (bc * H + 2 * yi) * W + 2 * xi;
        y[i] = 0.25f * (((p[0] + p[1]) + p[W]) + p[W + 1]);
    }
}

// dx[b,c,y,x] += 0.25 dy[b,c,y/2,x/2]
__global__ void k_pool2_backward_add(const float *__restrict__ dy, float *__restrict__ dx, int BC, int H, int W)
{
    const size_t n = size_t(BC) * H * W;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t bc = i / (size_t(H) * W);
        const int r = int(i % (size_t(H) * W)), yi = r / W, xi = r % W;
        dx[i] += 0.25f * dy[(bc * (H / 2) + yi / 2) * (W / 2) + xi / 2];
    }
}

// z = up2(lo) + skip
__global__ void k_upadd(const float *__restrict__ lo, const float *__restrict__ skip, float *__restrict__ z, int BC,
                        int H, int W)
{
    const size_t n = size_t(BC) * H * W;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t bc = i / (size_t(H) * W);
        const int r = int(i % (size_t(H) * W)), yi = r / W, xi = r % W;
        z[i] = lo[(bc * (H / 2) + yi / 2) * (W / 2) + xi / 2] + skip[i];
    }
}

// dlo[b,c,i,j] = sum of dz over the 2x2 block (dlo: H/2 x W/2)
__global__ void k_upsample_backward(const float *__restrict__ dz, float *__restrict__ dlo, int BC, int H, int W)
{
    const int h = H / 2, w = W / 2;
    const size_t n = size_t(BC) * h * w;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t bc = i / (size_t(h) * w);
        const int r = int(i % (size_t(h) * w)), yi = r / w, xi = r % w;
        const float *p = dz + (bc * H + 2 * yi) * W + 2 * xi;
        dlo[i] = ((p[0] + p[1]) + p[W]) + p[W + 1];
    }
}

// g *= [a > 0]
__global__ void k_relu_mask(float *__restrict__ g, const float *__restrict__ a, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        if (!(a[i] > 0.f))
            g[i] = 0.f;
}

constexpr int kLossBlocks = 256;

// dout = 2 m (out - t) norm; partial[block] = sum m (out - t)^2 (FP64)
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ out, const float *__restrict__ tgt,
                                              const float *__restrict__ in, float *__restrict__ dout,
                                              double *__restrict__ partial, int B, int HW, float norm)
{
    const size_t n = size_t(B) * LBM_SURROGATE_OUTPUTS * HW;
    double s = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t b = i / (size_t(LBM_SURROGATE_OUTPUTS) * HW);
        const size_t p = i % size_t(HW);
        const float m = in[b * LBM_SURROGATE_INPUTS * size_t(HW) + p];
        const float e = out[i] - tgt[i];
        s += double(m) * double(e) * double(e);
        dout[i] = 2.f * m * e * norm;
    }
    __shared__ double part[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = ((part[0] + part[1]) + part[2]) + part[3];
}

__global__ void k_loss_final(const double *__restrict__ partial, int blocks, double norm, double *__restrict__ loss,
                             float *__restrict__ history, int slot)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double s = 0.0;
    for (int i = 0; i < blocks; ++i)
        s += partial[i];
    s *= norm;
    *loss = s;
    if (history)
        history[slot] = float(s);
}

// Adam step (bias-corrected): c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, size_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

inline unsigned elem_grid(size_t n) { return unsigned(std::min<size_t>((n + 255) / 256, 8192)); }

uint64_t splitmix64(uint64_t &state)
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &state) { return double(splitmix64(state) >> 11) * 0x1.0p-53; }

#define S_TRY(expr)                                                                                    \
    do                                                                                                 \
    {                                                                                                  \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return lbm_set_error(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_DEVICE,   \
                                 std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

}  // namespace

struct lbm_surrogate
{
    int H = 0, W = 0, Bmax = 0;
    int LH[5] = {}, LW[5] = {};
    size_t nparams = 0;
    size_t w_off[kConvs] = {}, b_off[kConvs] = {};
    float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    float *wp[kConvs] = {}, *wt[kConvs] = {};
    short *wb[kConvs] = {};  // BF16 [Cout16][K16] weights of the matrix-core inference
    bool packed = false;
    uint32_t adam_t = 0;
    float *in = nullptr, *tgt = nullptr, *out = nullptr, *dout = nullptr, *dtmp = nullptr;
    float *a[14] = {}, *p[5] = {}, *z[4] = {}, *da[14] = {};
    float *partial = nullptr;
    size_t partial_floats = 0;
    double *part2 = nullptr;  // kReduceGroups x (weights + biases) of the largest conv
    double *loss_part = nullptr, *loss_dev = nullptr;
    float *history = nullptr;
    size_t history_cap = 0;
    float *ds_in = nullptr, *ds_tgt = nullptr;
    size_t ds_cap = 0, ds_n = 0;
    std::vector<double> ds_fluid;
    // training schedule and shuffle state, carried across train calls
    uint32_t sched_step = 0;
    bool shuffle_seeded = false;
    uint64_t shuffle_state = 0;
    std::vector<uint32_t> perm;
    size_t cursor = 0;
    std::vector<void *> allocations;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

namespace
{

size_t param_count()
{
    size_t n = 0;
    for (int k = 0; k < kConvs; ++k)
        n += size_t(kCout[k]) * kCin[k] * 9 + kCout[k];
    return n;
}

int alloc(lbm_surrogate *s, float **ptr, size_t floats)
{
    void *v = nullptr;
    S_TRY(hipMalloc(&v, std::max<size_t>(floats, 1) * sizeof(float)));
    s->allocations.push_back(v);
    *ptr = static_cast<float *>(v);
    return LBM_OK;
}

size_t act_floats(const lbm_surrogate *s, int C, int level)
{
    return size_t(s->Bmax) * C * s->LH[level] * s->LW[level];
}

int tiles_x(const lbm_surrogate *s, int l) { return (s->LW[l] + kTile - 1) / kTile; }
int tiles(const lbm_surrogate *s, int l) { return tiles_x(s, l) * ((s->LH[l] + kTile - 1) / kTile); }

// Weight-gradient slices: enough blocks to fill the GPU, at most one tile each.
int wgrad_slices(const lbm_surrogate *s, int k, int l, int B)
{
    const int blocks = (round16(kCout[k]) / kWgT) * (round16(kCin[k]) / kWgT);
    const int want = std::max(1, 2048 / blocks);
    return std::min(want, B * tiles(s, l));
}

int pack_all(lbm_surrogate *s)
{
    for (int k = 0; k < kConvs; ++k)
    {
        const int n = kCout[k] * kCin[k] * 9;
        hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->params + s->w_off[k], s->wp[k],
                           s->wt[k], kCin[k], kCout[k], round16(9 * kCin[k]), round16(9 * kCout[k]));
        hipLaunchKernelGGL(k_pack_bf16, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->params + s->w_off[k],
                           s->wb[k], kCin[k], kCout[k], round16(9 * kCin[k]));
    }
    S_TRY(hipGetLastError());
    s->packed = true;
    return LBM_OK;
}

// FP32 convolution of x (Cin channels) with packed weights wm ([Cout16][round16(9 Cin)]) into y.
void conv_f32(lbm_surrogate *s, const float *x, const float *wm, const float *bias, float *y, int Cin, int Cout, int l,
              int B, bool relu)
{
    const int tx = (s->LW[l] + kMfmaCols - 1) / kMfmaCols, ty = (s->LH[l] + kMfmaRows - 1) / kMfmaRows;
    const size_t lds = size_t(Cin) * (kMfmaRows + 2) * (kMfmaCols + 2) * sizeof(float);
    hipLaunchKernelGGL(k_conv3x3_f32, dim3(unsigned(tx * ty), unsigned(B)), dim3(256), lds, s->stream, x, wm, bias, y,
                       Cin, Cout, round16(9 * Cin), s->LH[l], s->LW[l], tx, relu ? 1 : 0);
}

void conv_forward_bf16(lbm_surrogate *s, const float *x, int k, float *y, int l, int B, bool relu)
{
    const int tx = (s->LW[l] + kMfmaCols - 1) / kMfmaCols, ty = (s->LH[l] + kMfmaRows - 1) / kMfmaRows;
    const size_t lds = size_t(kCin[k]) * (kMfmaRows + 2) * (kMfmaCols + 2) * sizeof(short);
    hipLaunchKernelGGL(k_conv3x3_bf16, dim3(unsigned(tx * ty), unsigned(B)), dim3(256), lds, s->stream, x, s->wb[k],
                       s->params + s->b_off[k], y, kCin[k], kCout[k], round16(9 * kCin[k]), s->LH[l], s->LW[l], tx,
                       relu ? 1 : 0);
}

void conv_forward(lbm_surrogate *s, const float *x, int k, float *y, int l, int B, bool relu, bool bf16 = false)
{
    if (bf16)
    {
        conv_forward_bf16(s, x, k, y, l, B, relu);
        return;
    }
    conv_f32(s, x, s->wp[k], s->params + s->b_off[k], y, kCin[k], kCout[k], l, B, relu);
}

// dx = data gradient of conv k for output gradient g (dx written, not accumulated)
void conv_dgrad(lbm_surrogate *s, const float *g, int k, float *dx, int l, int B)
{
    conv_f32(s, g, s->wt[k], nullptr, dx, kCout[k], kCin[k], l, B, false);
}

void conv_wgrad(lbm_surrogate *s, const float *x, const float *g, int k, int l, int B)
{
    const int CinP = round16(kCin[k]), CoutP = round16(kCout[k]);
    const int slices = wgrad_slices(s, k, l, B);
    const dim3 grid(unsigned((CoutP / kWgT) * (CinP / kWgT)), unsigned(slices));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, s->stream, x, g, s->partial, B, kCin[k], kCout[k], CinP,
                       CoutP, s->LH[l], s->LW[l], tiles_x(s, l), tiles(s, l), slices);
    const int n = kCout[k] * kCin[k] * 9 + kCout[k];
    const int groups = std::min(kReduceGroups, slices), per = (slices + groups - 1) / groups;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((n + 255) / 256, groups), dim3(256), 0, s->stream, s->partial,
                       s->part2, kCin[k], kCout[k], CinP, CoutP, slices, per);
    hipLaunchKernelGGL(k_wgrad_final, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->part2,
                       s->grad + s->w_off[k], s->grad + s->b_off[k], kCin[k], kCout[k], groups);
}

void pool(lbm_surrogate *s, const float *x, float *y, int C, int l_in, int B)
{
    const size_t n = size_t(B) * C * s->LH[l_in + 1] * s->LW[l_in + 1];
    hipLaunchKernelGGL(k_pool2, dim3(elem_grid(n)), dim3(256), 0, s->stream, x, y, B * C, s->LH[l_in], s->LW[l_in]);
}

void pool_backward_add(lbm_surrogate *s, const float *dy, float *dx, int C, int l_in, int B)
{
    const size_t n = size_t(B) * C * s->LH[l_in] * s->LW[l_in];
    hipLaunchKernelGGL(k_pool2_backward_add, dim3(elem_grid(n)), dim3(256), 0, s->stream, dy, dx, B * C,
                       s->LH[l_in], s->LW[l_in]);
}

void upadd(lbm_surrogate *s, const float *lo, const float *skip, float *z, int C, int l, int B)
{
    const size_t n = size_t(B) * C * s->LH[l] * s->LW[l];
    hipLaunchKernelGGL(k_upadd, dim3(elem_grid(n)), dim3(256), 0, s->stream, lo, skip, z, B * C, s->LH[l], s->LW[l]);
}

void upsample_backward(lbm_surrogate *s, const float *dz, float *dlo, int C, int l, int B)
{
    const size_t n = size_t(B) * C * s->LH[l + 1] * s->LW[l + 1];
    hipLaunchKernelGGL(k_upsample_backward, dim3(elem_grid(n)), dim3(256), 0, s->stream, dz, dlo, B * C, s->LH[l],
                       s->LW[l]);
}

void relu_mask(lbm_surrogate *s, float *g, const float *a, int C, int l, int B)
{
    const size_t n = size_t(B) * C * s->LH[l] * s->LW[l];
    hipLaunchKernelGGL(k_relu_mask, dim3(elem_grid(n)), dim3(256), 0, s->stream, g, a, n);
}

// Forward over s->in (B samples) into s->a, s->p, s->z, s->out; bf16: the
// convolutions on the matrix cores (inference).
void forward(lbm_surrogate *s, int B, bool bf16 = false)
{
    conv_forward(s, s->in, 0, s->a[1], 0, B, true, bf16);
    conv_forward(s, s->a[1], 1, s->a[2], 0, B, true, bf16);
    for (int l = 1; l <= 3; ++l)
    {
        pool(s, s->a[2 * l], s->p[l], kCout[2 * l - 1], l - 1, B);
        conv_forward(s, s->p[l], 2 * l, s->a[2 * l + 1], l, B, true, bf16);
        conv_forward(s, s->a[2 * l + 1], 2 * l + 1, s->a[2 * l + 2], l, B, true, bf16);
    }
    pool(s, s->a[8], s->p[4], kCout[7], 3, B);
    conv_forward(s, s->p[4], 8, s->a[9], 4, B, true, bf16);
    for (int m = 0; m < 4; ++m)
    {
        const int l = 3 - m, kk = 9 + m;
        upadd(s, s->a[9 + m], s->a[8 - 2 * m], s->z[l], kCin[kk], l, B);
        conv_forward(s, s->z[l], kk, s->a[10 + m], l, B, true, bf16);
    }
    conv_forward(s, s->a[13], 13, s->out, 0, B, false, bf16);
}

// Loss of s->out against s->tgt and the full parameter gradient into s->grad.
void loss_backward(lbm_surrogate *s, int B, double norm, float *history, int slot)
{
    const int HW = s->H * s->W;
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, s->stream, s->out, s->tgt, s->in, s->dout,
                       s->loss_part, B, HW, float(norm));
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, s->stream, s->loss_part, kLossBlocks, norm, s->loss_dev,
                       history, slot);
    // conv13 (no ReLU)
    conv_wgrad(s, s->a[13], s->dout, 13, 0, B);
    conv_dgrad(s, s->dout, 13, s->da[13], 0, B);
    for (int m = 3; m >= 0; --m)
    {
        const int l = 3 - m, kk = 9 + m;
        relu_mask(s, s->da[10 + m], s->a[10 + m], kCout[kk], l, B);
        conv_wgrad(s, s->z[l], s->da[10 + m], kk, l, B);
        // dz is the skip's gradient: written straight into da[8-2m]
        conv_dgrad(s, s->da[10 + m], kk, s->da[8 - 2 * m], l, B);
        upsample_backward(s, s->da[8 - 2 * m], s->da[9 + m], kCin[kk], l, B);
    }
    relu_mask(s, s->da[9], s->a[9], kCout[8], 4, B);
    conv_wgrad(s, s->p[4], s->da[9], 8, 4, B);
    conv_dgrad(s, s->da[9], 8, s->dtmp, 4, B);
    pool_backward_add(s, s->dtmp, s->da[8], kCin[8], 3, B);
    for (int l = 3; l >= 0; --l)
    {
        const int k1 = 2 * l + 1, k0 = 2 * l;
        relu_mask(s, s->da[k1 + 1], s->a[k1 + 1], kCout[k1], l, B);
        conv_wgrad(s, s->a[k1], s->da[k1 + 1], k1, l, B);
        conv_dgrad(s, s->da[k1 + 1], k1, s->da[k1], l, B);
        relu_mask(s, s->da[k1], s->a[k1], kCout[k0], l, B);
        conv_wgrad(s, l == 0 ? s->in : s->p[l], s->da[k1], k0, l, B);
        if (l > 0)
        {
            conv_dgrad(s, s->da[k1], k0, s->dtmp, l, B);
            pool_backward_add(s, s->dtmp, s->da[k0], kCin[k0], l - 1, B);
        }
    }
}

int adam_step(lbm_surrogate *s, float lr)
{
    const double b1 = 0.9, b2 = 0.999;
    ++s->adam_t;
    const double c1 = 1.0 - std::pow(b1, double(s->adam_t)), c2 = 1.0 - std::pow(b2, double(s->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(elem_grid(s->nparams)), dim3(256), 0, s->stream, s->params, s->adam_m, s->adam_v,
                       s->grad, s->nparams, lr, float(b1), float(b2), 1e-8f, float(c1), float(c2));
    S_TRY(hipGetLastError());
    return pack_all(s);
}

int reset_adam(lbm_surrogate *s)
{
    s->sched_step = 0;
    s->shuffle_seeded = false;
    S_TRY(hipMemsetAsync(s->adam_m, 0, s->nparams * sizeof(float), s->stream));
    S_TRY(hipMemsetAsync(s->adam_v, 0, s->nparams * sizeof(float), s->stream));
    s->adam_t = 0;
    return LBM_OK;
}

double fluid_sum(const float *inputs, uint32_t batch, size_t hw)
{
    double m = 0.0;
    for (uint32_t b = 0; b < batch; ++b)
    {
        const float *c0 = inputs + size_t(b) * LBM_SURROGATE_INPUTS * hw;
        for (size_t i = 0; i < hw; ++i)
            m += double(c0[i]);
    }
    return m;
}

int check_ready(const lbm_surrogate *s, const char *what)
{
    if (!s)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, std::string(what) + ": null handle");
    if (!s->packed)
        return lbm_set_error(LBM_ERR_STATE, std::string(what) + ": parameters not set (init or set them first)");
    return LBM_OK;
}

}  // namespace

extern "C" {

size_t lbm_surrogate_param_count(void) { return param_count(); }

void lbm_surrogate_destroy(lbm_surrogate *s)
{
    if (!s)
        return;
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    for (void *v : s->allocations)
        (void)hipFree(v);
    (void)hipFree(s->loss_part);
    (void)hipFree(s->loss_dev);
    (void)hipFree(s->history);
    (void)hipFree(s->ds_in);
    (void)hipFree(s->ds_tgt);
    if (s->ev0)
        (void)hipEventDestroy(s->ev0);
    if (s->ev1)
        (void)hipEventDestroy(s->ev1);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

int lbm_surrogate_create(uint32_t height, uint32_t width, uint32_t max_batch, lbm_surrogate **out)
{
    if (!out)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_create: null output");
    *out = nullptr;
    if (height == 0 || width == 0 || height % 16 || width % 16 || height > 8192 || width > 8192 || max_batch == 0 ||
        max_batch > 4096)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT,
                             "lbm_surrogate_create: height and width must be positive multiples of 16 (<= 8192), "
                             "max_batch in 1..4096");
    auto *s = new lbm_surrogate;
    s->H = int(height);
    s->W = int(width);
    s->Bmax = int(max_batch);
    for (int l = 0; l < 5; ++l)
    {
        s->LH[l] = s->H >> l;
        s->LW[l] = s->W >> l;
    }
    size_t off = 0;
    for (int k = 0; k < kConvs; ++k)
    {
        s->w_off[k] = off;
        off += size_t(kCout[k]) * kCin[k] * 9;
        s->b_off[k] = off;
        off += kCout[k];
    }
    s->nparams = off;
    int rc = LBM_OK;
#define C_TRY(expr)                                                                                    \
    do                                                                                                 \
    {                                                                                                  \
        if ((rc = (expr)) != LBM_OK)                                                                   \
        {                                                                                              \
            lbm_surrogate_destroy(s);                                                                  \
            return rc;                                                                                 \
        }                                                                                              \
    } while (0)
    C_TRY([&]() -> int {
        S_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        S_TRY(hipEventCreate(&s->ev0));
        S_TRY(hipEventCreate(&s->ev1));
        S_TRY(hipMalloc(&s->loss_part, kLossBlocks * sizeof(double)));
        S_TRY(hipMalloc(&s->loss_dev, sizeof(double)));
        return LBM_OK;
    }());
    C_TRY(alloc(s, &s->params, s->nparams));
    C_TRY(alloc(s, &s->grad, s->nparams));
    C_TRY(alloc(s, &s->adam_m, s->nparams));
    C_TRY(alloc(s, &s->adam_v, s->nparams));
    size_t partial = 0;
    for (int k = 0; k < kConvs; ++k)
    {
        const size_t cin_p = size_t(round16(kCin[k])), cout_p = size_t(round16(kCout[k]));
        C_TRY(alloc(s, &s->wp[k], cin_p * 9 * cout_p));
        C_TRY(alloc(s, &s->wt[k], cin_p * 9 * cout_p));
        const size_t kp = size_t(round16(9 * kCin[k]));
        float *wbf = nullptr;
        C_TRY(alloc(s, &wbf, (cout_p * kp + 1) / 2));
        s->wb[k] = reinterpret_cast<short *>(wbf);
        C_TRY([&]() -> int {
            S_TRY(hipMemsetAsync(s->wp[k], 0, cin_p * 9 * cout_p * sizeof(float), s->stream));
            S_TRY(hipMemsetAsync(s->wt[k], 0, cin_p * 9 * cout_p * sizeof(float), s->stream));
            S_TRY(hipMemsetAsync(s->wb[k], 0, cout_p * kp * sizeof(short), s->stream));
            return LBM_OK;
        }());
        partial = std::max(partial, size_t(wgrad_slices(s, k, conv_level(k), s->Bmax)) * cout_p * cin_p * 10);
    }
    s->partial_floats = partial;
    C_TRY(alloc(s, &s->partial, partial));
    {
        size_t most = 0;
        for (int k = 0; k < kConvs; ++k)
            most = std::max(most, size_t(kCout[k]) * kCin[k] * 9 + kCout[k]);
        float *p2 = nullptr;
        C_TRY(alloc(s, &p2, 2 * kReduceGroups * most));  // doubles
        s->part2 = reinterpret_cast<double *>(p2);
    }
    C_TRY(alloc(s, &s->in, act_floats(s, LBM_SURROGATE_INPUTS, 0)));
    C_TRY(alloc(s, &s->tgt, act_floats(s, LBM_SURROGATE_OUTPUTS, 0)));
    C_TRY(alloc(s, &s->out, act_floats(s, LBM_SURROGATE_OUTPUTS, 0)));
    C_TRY(alloc(s, &s->dout, act_floats(s, LBM_SURROGATE_OUTPUTS, 0)));
    // a[k+1] = output of conv k (k < 9) at level k/2 (conv8: level 4); decoder a[10+m] at level 3-m
    for (int k = 0; k < 9; ++k)
    {
        const int level = k < 8 ? k / 2 : 4;
        C_TRY(alloc(s, &s->a[k + 1], act_floats(s, kCout[k], level)));
        C_TRY(alloc(s, &s->da[k + 1], act_floats(s, kCout[k], level)));
    }
    for (int m = 0; m < 4; ++m)
    {
        C_TRY(alloc(s, &s->a[10 + m], act_floats(s, kCout[9 + m], 3 - m)));
        C_TRY(alloc(s, &s->da[10 + m], act_floats(s, kCout[9 + m], 3 - m)));
        C_TRY(alloc(s, &s->z[3 - m], act_floats(s, kCin[9 + m], 3 - m)));
    }
    for (int l = 1; l <= 4; ++l)
        C_TRY(alloc(s, &s->p[l], act_floats(s, kCout[2 * l - 1], l)));
    size_t tmp = 0;
    for (int l = 1; l <= 4; ++l)
        tmp = std::max(tmp, act_floats(s, kCin[l == 4 ? 8 : 2 * l], l));
    C_TRY(alloc(s, &s->dtmp, tmp));
    C_TRY([&]() -> int {
        S_TRY(hipStreamSynchronize(s->stream));
        return LBM_OK;
    }());
#undef C_TRY
    *out = s;
    return LBM_OK;
}

int lbm_surrogate_init_params(lbm_surrogate *s, uint64_t seed)
{
    if (!s)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_init_params: null handle");
    std::vector<float> host(s->nparams, 0.f);
    uint64_t st = seed;
    for (int k = 0; k < kConvs; ++k)
    {
        const double stdev = std::sqrt(2.0 / (9.0 * kCin[k]));
        const size_t n = size_t(kCout[k]) * kCin[k] * 9;
        for (size_t i = 0; i < n; i += 2)
        {
            // Box-Muller pair
            const double u1 = std::max(uniform01(st), 1e-300), u2 = uniform01(st);
            const double r = std::sqrt(-2.0 * std::log(u1)), t = 2.0 * kPi * u2;
            host[s->w_off[k] + i] = float(stdev * r * std::cos(t));
            if (i + 1 < n)
                host[s->w_off[k] + i + 1] = float(stdev * r * std::sin(t));
        }
    }
    return lbm_surrogate_set_params(s, host.data(), host.size());
}

int lbm_surrogate_set_params(lbm_surrogate *s, const float *params, size_t count)
{
    if (!s || !params || count != s->nparams)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_set_params: count must equal the parameter count");
    S_TRY(hipMemcpyAsync(s->params, params, count * sizeof(float), hipMemcpyHostToDevice, s->stream));
    if (int rc = reset_adam(s))
        return rc;
    if (int rc = pack_all(s))
        return rc;
    S_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

int lbm_surrogate_get_params(lbm_surrogate *s, float *params, size_t count)
{
    if (int rc = check_ready(s, "lbm_surrogate_get_params"))
        return rc;
    if (!params || count != s->nparams)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_get_params: count must equal the parameter count");
    S_TRY(hipMemcpyAsync(params, s->params, count * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

static int predict(lbm_surrogate *s, const float *inputs, uint32_t batch, float *outputs, bool bf16)
{
    if (int rc = check_ready(s, "lbm_surrogate_predict"))
        return rc;
    if (!inputs || !outputs)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_predict: null buffer");
    const size_t hw = size_t(s->H) * s->W;
    S_TRY(hipEventRecord(s->ev0, s->stream));
    for (uint32_t b0 = 0; b0 < batch; b0 += uint32_t(s->Bmax))
    {
        const int B = int(std::min<uint32_t>(uint32_t(s->Bmax), batch - b0));
        S_TRY(hipMemcpyAsync(s->in, inputs + size_t(b0) * LBM_SURROGATE_INPUTS * hw,
                             size_t(B) * LBM_SURROGATE_INPUTS * hw * sizeof(float), hipMemcpyHostToDevice, s->stream));
        forward(s, B, bf16);
        S_TRY(hipGetLastError());
        S_TRY(hipMemcpyAsync(outputs + size_t(b0) * LBM_SURROGATE_OUTPUTS * hw, s->out,
                             size_t(B) * LBM_SURROGATE_OUTPUTS * hw * sizeof(float), hipMemcpyDeviceToHost,
                             s->stream));
    }
    S_TRY(hipEventRecord(s->ev1, s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    S_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    return LBM_OK;
}

int lbm_surrogate_predict(lbm_surrogate *s, const float *inputs, uint32_t batch, float *outputs)
{
    return predict(s, inputs, batch, outputs, false);
}

int lbm_surrogate_predict_bf16(lbm_surrogate *s, const float *inputs, uint32_t batch, float *outputs)
{
    return predict(s, inputs, batch, outputs, true);
}

static int host_batch(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t batch, double *norm,
                      const char *what)
{
    if (int rc = check_ready(s, what))
        return rc;
    if (!inputs || !targets || batch == 0 || batch > uint32_t(s->Bmax))
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, std::string(what) + ": batch must be in 1..max_batch");
    const size_t hw = size_t(s->H) * s->W;
    const double m = fluid_sum(inputs, batch, hw);
    if (!(m > 0.0))
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, std::string(what) + ": the batch has no fluid node");
    *norm = 1.0 / (LBM_SURROGATE_OUTPUTS * m);
    S_TRY(hipMemcpyAsync(s->in, inputs, size_t(batch) * LBM_SURROGATE_INPUTS * hw * sizeof(float),
                         hipMemcpyHostToDevice, s->stream));
    S_TRY(hipMemcpyAsync(s->tgt, targets, size_t(batch) * LBM_SURROGATE_OUTPUTS * hw * sizeof(float),
                         hipMemcpyHostToDevice, s->stream));
    return LBM_OK;
}

int lbm_surrogate_loss_grad(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t batch,
                            double *loss, float *grad, size_t count)
{
    double norm = 0.0;
    if (int rc = host_batch(s, inputs, targets, batch, &norm, "lbm_surrogate_loss_grad"))
        return rc;
    if (!loss || !grad || count != s->nparams)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_loss_grad: bad output buffers");
    forward(s, int(batch));
    loss_backward(s, int(batch), norm, nullptr, 0);
    S_TRY(hipGetLastError());
    S_TRY(hipMemcpyAsync(loss, s->loss_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    S_TRY(hipMemcpyAsync(grad, s->grad, count * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

int lbm_surrogate_train_step(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t batch,
                             float learning_rate, double *loss)
{
    double norm = 0.0;
    if (int rc = host_batch(s, inputs, targets, batch, &norm, "lbm_surrogate_train_step"))
        return rc;
    forward(s, int(batch));
    loss_backward(s, int(batch), norm, nullptr, 0);
    S_TRY(hipGetLastError());
    if (int rc = adam_step(s, learning_rate))
        return rc;
    if (loss)
        S_TRY(hipMemcpyAsync(loss, s->loss_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

int lbm_surrogate_dataset_reserve(lbm_surrogate *s, size_t capacity)
{
    if (!s || capacity == 0)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_dataset_reserve: bad argument");
    const size_t hw = size_t(s->H) * s->W;
    (void)hipFree(s->ds_in);
    (void)hipFree(s->ds_tgt);
    s->ds_in = s->ds_tgt = nullptr;
    s->ds_cap = s->ds_n = 0;
    s->ds_fluid.clear();
    s->shuffle_seeded = false;
    S_TRY(hipMalloc(&s->ds_in, capacity * LBM_SURROGATE_INPUTS * hw * sizeof(float)));
    S_TRY(hipMalloc(&s->ds_tgt, capacity * LBM_SURROGATE_OUTPUTS * hw * sizeof(float)));
    s->ds_cap = capacity;
    return LBM_OK;
}

int lbm_surrogate_dataset_add(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t count)
{
    if (!s || !inputs || !targets)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_dataset_add: bad argument");
    if (s->ds_n + count > s->ds_cap)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_dataset_add: capacity exceeded");
    const size_t hw = size_t(s->H) * s->W;
    for (uint32_t i = 0; i < count; ++i)
        s->ds_fluid.push_back(fluid_sum(inputs + size_t(i) * LBM_SURROGATE_INPUTS * hw, 1, hw));
    S_TRY(hipMemcpyAsync(s->ds_in + s->ds_n * LBM_SURROGATE_INPUTS * hw, inputs,
                         size_t(count) * LBM_SURROGATE_INPUTS * hw * sizeof(float), hipMemcpyHostToDevice, s->stream));
    S_TRY(hipMemcpyAsync(s->ds_tgt + s->ds_n * LBM_SURROGATE_OUTPUTS * hw, targets,
                         size_t(count) * LBM_SURROGATE_OUTPUTS * hw * sizeof(float), hipMemcpyHostToDevice,
                         s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    s->ds_n += count;
    s->shuffle_seeded = false;
    return LBM_OK;
}

size_t lbm_surrogate_dataset_size(const lbm_surrogate *s) { return s ? s->ds_n : 0; }

int lbm_surrogate_train(lbm_surrogate *s, uint32_t steps, uint32_t schedule_steps, uint32_t batch, float lr_max,
                        float lr_min, uint64_t seed, float *loss_history)
{
    if (int rc = check_ready(s, "lbm_surrogate_train"))
        return rc;
    if (batch == 0 || batch > uint32_t(s->Bmax) || s->ds_n < batch || schedule_steps == 0)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT,
                             "lbm_surrogate_train: batch must be in 1..max_batch and at most the dataset size, "
                             "schedule_steps > 0");
    if (steps == 0)
        return LBM_OK;
    if (s->history_cap < steps)
    {
        (void)hipFree(s->history);
        s->history = nullptr;
        s->history_cap = 0;
        S_TRY(hipMalloc(&s->history, size_t(steps) * sizeof(float)));
        s->history_cap = steps;
    }
    const size_t hw = size_t(s->H) * s->W;
    if (!s->shuffle_seeded)
    {
        s->perm.resize(s->ds_n);
        for (size_t i = 0; i < s->perm.size(); ++i)
            s->perm[i] = uint32_t(i);
        s->shuffle_state = seed;
        s->cursor = s->perm.size();
        s->shuffle_seeded = true;
    }
    std::vector<uint32_t> &perm = s->perm;
    const uint32_t warm = std::max<uint32_t>(1, std::min<uint32_t>(200, schedule_steps / 10));
    S_TRY(hipEventRecord(s->ev0, s->stream));
    for (uint32_t step = 0; step < steps; ++step)
    {
        if (s->cursor + batch > perm.size())
        {
            for (size_t i = perm.size() - 1; i > 0; --i)
                std::swap(perm[i], perm[splitmix64(s->shuffle_state) % (i + 1)]);
            s->cursor = 0;
        }
        const size_t cursor = s->cursor;
        double m = 0.0;
        for (uint32_t i = 0; i < batch; ++i)
        {
            const size_t id = perm[cursor + i];
            m += s->ds_fluid[id];
            S_TRY(hipMemcpyAsync(s->in + size_t(i) * LBM_SURROGATE_INPUTS * hw,
                                 s->ds_in + id * LBM_SURROGATE_INPUTS * hw, LBM_SURROGATE_INPUTS * hw * sizeof(float),
                                 hipMemcpyDeviceToDevice, s->stream));
            S_TRY(hipMemcpyAsync(s->tgt + size_t(i) * LBM_SURROGATE_OUTPUTS * hw,
                                 s->ds_tgt + id * LBM_SURROGATE_OUTPUTS * hw,
                                 LBM_SURROGATE_OUTPUTS * hw * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        }
        s->cursor += batch;
        if (!(m > 0.0))
            return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_train: a minibatch has no fluid node");
        const uint32_t g = s->sched_step++;
        double lr;
        if (g < warm)
            lr = lr_max * double(g + 1) / warm;
        else
        {
            const double t = std::min(1.0, double(g - warm) / std::max<uint32_t>(1, schedule_steps - warm));
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + std::cos(kPi * t));
        }
        forward(s, int(batch));
        loss_backward(s, int(batch), 1.0 / (LBM_SURROGATE_OUTPUTS * m), s->history, int(step));
        S_TRY(hipGetLastError());
        if (int rc = adam_step(s, float(lr)))
            return rc;
    }
    S_TRY(hipEventRecord(s->ev1, s->stream));
    if (loss_history)
        S_TRY(hipMemcpyAsync(loss_history, s->history, size_t(steps) * sizeof(float), hipMemcpyDeviceToHost,
                             s->stream));
    S_TRY(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    S_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    return LBM_OK;
}

int lbm_surrogate_last_ms(const lbm_surrogate *s, double *ms)
{
    if (!s || !ms)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_surrogate_last_ms: bad argument");
    *ms = s->last_ms;
    return LBM_OK;
}

}  // extern "C"
