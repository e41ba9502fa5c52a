// prior.hip -- gfx950 kernels of the learned 2.5D residual prior (no reference counterpart).
//
// 3x3 "same" convolutions over NHWC bfloat16 activations with bfloat16 weights [Cout][3][3][Cin]
// and FP32 accumulation:
//   conv3x3_fwd_kernel    implicit GEMM on MFMA (v_mfma_f32_16x16x32_bf16): M = pixels,
//                         N = Cout, K = 9*Cin; epilogue adds bias, optional ReLU, optional
//                         residual, stores bfloat16.  The input gradient of a convolution is
//                         the same kernel applied to the output gradient with the weights
//                         flipped in space and transposed in channels.
//   conv3x3_fwd_small_cin_kernel  the same for at most 4 input channels, input halo in LDS
//   conv3x3_fwd_persistent_kernel  the same for 64 input channels and 64 or at most 16 output
//                         channels: one workgroup per CU with all weights in LDS, walking pixel
//                         tiles with the next tile's input halo loaded while the current one is
//                         computed.
//   conv3x3_wgrad_kernel  weight and bias gradients on MFMA: per (pixel chunk, tap), 64-pixel
//                         tiles of the output gradient and of the shifted input staged in LDS
//                         transposed to [channel][pixel]; grad_w = gout^T x_shift with K =
//                         pixels, FP32 accumulation; partial sums per chunk
//   conv3x3_wgrad_reduce  sums the chunk partials in a fixed order (deterministic)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ctrecon.h"

namespace {

typedef __bf16 bf16_t;
typedef bf16_t bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16_t bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kConvWaves = 4;    // waves per workgroup
constexpr int kPxPerWave = 32;   // two 16-pixel MFMA row tiles per wave
constexpr int kMaxCout = 64;     // four 16-channel MFMA column tiles

__device__ inline bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e) z[e] = (bf16_t)0.0f;
    return z;
}

// MFMA operand layout (16 x 16 x 32, wave64): lane l supplies row (l % 16) of A and column
// (l % 16) of B, both with the 8 consecutive K elements 8*(l / 16) .. 8*(l / 16) + 7; it receives
// rows 4*(l / 16) .. +3 of column (l % 16) of D.
template <bool FAST>
__global__ void __launch_bounds__(256) conv3x3_fwd_kernel(
    const bf16_t* __restrict__ in, const bf16_t* __restrict__ w, const float* __restrict__ bias,
    const bf16_t* __restrict__ resid, bf16_t* __restrict__ out, int N, int H, int W, int Cin,
    int Cout, int relu, int vec) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = N * H * W;
    const int pbase = (blockIdx.x * kConvWaves + wave) * kPxPerWave;
    const int m = lane & 15, g = lane >> 4;
    const int K = 9 * Cin;
    const int ksteps = (K + 31) / 32;
    const int n_tiles = (Cout + 15) / 16;
    int py[2], px[2], pn[2];
    bool pv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = pbase + h * 16 + m;
        pv[h] = p < P;
        const int pp = pv[h] ? p : 0;
        px[h] = pp % W;
        const int t = pp / W;
        py[h] = t % H;
        pn[h] = t / H;
    }
    f32x4 acc[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int s = 0; s < ksteps; ++s) {
        const int k0 = s * 32 + g * 8;
        bf16x8 a[2], b[4];
        if constexpr (FAST) {
            // Cin % 32 == 0: the 8 K elements are 8 channels of one tap, 16-byte aligned
            const int tap = k0 / Cin, ci = k0 - tap * Cin;
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int yy = py[h] + dy, xx = px[h] + dx;
                const bool ok = pv[h] && yy >= 0 && yy < H && xx >= 0 && xx < W;
                a[h] = ok ? *reinterpret_cast<const bf16x8*>(
                                in + ((size_t)(pn[h] * H + yy) * W + xx) * Cin + ci)
                          : zero8();
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int co = t * 16 + m;
                b[t] = (co < Cout) ? *reinterpret_cast<const bf16x8*>(w + (size_t)co * K + k0) : zero8();
            }
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) a[h] = zero8();
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = zero8();
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + e;
                if (k >= K) break;
                const int tap = k / Cin, ci = k - tap * Cin;
                const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int yy = py[h] + dy, xx = px[h] + dx;
                    if (pv[h] && yy >= 0 && yy < H && xx >= 0 && xx < W)
                        a[h][e] = in[((size_t)(pn[h] * H + yy) * W + xx) * Cin + ci];
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int co = t * 16 + m;
                    if (co < Cout) b[t][e] = w[(size_t)co * K + k];
                }
            }
        }
        // the weights are the row operand: lane (m, g) receives output channels
        // 16 t + 4 g .. +3 of pixel m, four adjacent values of one NHWC row
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < n_tiles)  // uniform: column tiles past Cout are all zero
                    acc[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[t], a[h], acc[h][t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (t >= n_tiles) continue;
        const int co0 = t * 16 + 4 * g;
        if (co0 >= Cout) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p = pbase + h * 16 + m;
            if (p >= P) continue;
            const size_t o = (size_t)p * Cout + co0;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[h][t][r] + ((bias && co0 + r < Cout) ? bias[co0 + r] : 0.0f);
                if (relu) v[r] = fmaxf(v[r], 0.0f);
            }
            if (vec) {  // Cout % 4 == 0 and 8-byte aligned rows: four channels per access
                if (resid) {
                    const bf16x4 rv = *reinterpret_cast<const bf16x4*>(resid + o);
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] += (float)rv[r];
                }
                bf16x4 q;
#pragma unroll
                for (int r = 0; r < 4; ++r) q[r] = (bf16_t)v[r];
                *reinterpret_cast<bf16x4*>(out + o) = q;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (co0 + r >= Cout) break;
                    float x = v[r];
                    if (resid) x += (float)resid[o + r];
                    out[o + r] = (bf16_t)x;
                }
            }
        }
    }
}

// Convolutions with at most 4 input channels (the network's first layer, 3 -> 64, and the input
// gradient of its last, 1 -> 64).  K = 9 taps x 4 channels (zero past Cin) in two MFMA K steps:
// taps 0-7, then tap 8.  A workgroup computes an 8 x 32 pixel tile from its (8+2) x (32+2) input
// halo in LDS (one 8-byte entry of 4 channels per pixel); lane (m, g) of wave w takes pixel m of
// the four 16-pixel row tiles of rows 2w, 2w+1 with K elements 8g .. 8g+7 (taps 2g and 2g+1 in
// the first step).  The workgroups are persistent: each builds the zero-padded weight operands
// in registers once and then walks tiles.
constexpr int kSTileH = 8, kSTileW = 32;
constexpr int kSHaloW = kSTileW + 2;
constexpr int kSHaloPix = (kSTileH + 2) * kSHaloW;

__global__ void __launch_bounds__(256) conv3x3_fwd_small_cin_kernel(
    const bf16_t* __restrict__ in, const bf16_t* __restrict__ w, const float* __restrict__ bias,
    const bf16_t* __restrict__ resid, bf16_t* __restrict__ out, int H, int W, int Cin, int Cout,
    int relu, int vec, int tiles_x, int tiles_y, int n_tiles) {
    __shared__ __attribute__((aligned(16))) bf16x4 halo[kSHaloPix];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int n_ct = (Cout + 15) / 16;  // column tiles (uniform)
    // weights as the row operand: lane (m, g) of column tile t holds output channel 16 t + m,
    // K elements k = 32 s + 8 g + e of step s (tap k / 4, channel k % 4)
    bf16x8 b[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 32 * s + 8 * g + e, tap = k >> 2, ci = k & 3, co = 16 * t + m;
                b[s][t][e] = (tap < 9 && ci < Cin && co < Cout) ? w[((size_t)co * 9 + tap) * Cin + ci] : (bf16_t)0.0f;
            }
    // lane (m, g) receives channels 16 t + 4 g .. +3 of its pixel: the bias starts them
    float bias_t[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = 16 * t + 4 * g + r;
            bias_t[t][r] = (bias && co < Cout) ? bias[co] : 0.0f;
        }
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, n = tile / (tiles_x * tiles_y);
        const int ty0 = ty * kSTileH, tx0 = tx * kSTileW;
        __syncthreads();  // the previous tile's halo is read
        for (int e = tid; e < kSHaloPix; e += 256) {
            const int yy = ty0 + e / kSHaloW - 1, xx = tx0 + e % kSHaloW - 1;
            bf16x4 v;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = (bf16_t)0.0f;
            if ((unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W) {
                const bf16_t* src = in + ((size_t)(n * H + yy) * W + xx) * Cin;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (c < Cin) v[c] = src[c];
            }
            halo[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 2 * wave + (i >> 1), col = 16 * (i & 1) + m;
            // halo entry of tap (dy, dx) for this pixel: (row + dy) * kSHaloW + col + dx
            const int t0 = 2 * g, t1 = 2 * g + 1;
            const bf16x4 p0 = halo[(row + t0 / 3) * kSHaloW + col + t0 % 3];
            const bf16x4 p1 = halo[(row + t1 / 3) * kSHaloW + col + t1 % 3];
            const bf16x4 p8 = halo[(row + 2) * kSHaloW + col + 2];
            bf16x8 a0, a1;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                a0[e] = p0[e];
                a0[4 + e] = p1[e];
                a1[e] = g == 0 ? p8[e] : (bf16_t)0.0f;
                a1[4 + e] = (bf16_t)0.0f;
            }
            const int yy = ty0 + row, xx = tx0 + col;
            const bool ok = yy < H && xx < W;
            const size_t o = ok ? ((size_t)(n * H + yy) * W + xx) * Cout : 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t >= n_ct) break;
                f32x4 acc = f32x4{bias_t[t][0], bias_t[t][1], bias_t[t][2], bias_t[t][3]};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0][t], a0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[1][t], a1, acc, 0, 0, 0);
                const int co0 = 16 * t + 4 * g;
                if (!ok || co0 >= Cout) continue;
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = relu ? fmaxf(acc[r], 0.0f) : acc[r];
                if (vec) {  // Cout % 4 == 0 and 8-byte aligned rows
                    if (resid) {
                        const bf16x4 rv = *reinterpret_cast<const bf16x4*>(resid + o + co0);
#pragma unroll
                        for (int r = 0; r < 4; ++r) v[r] += (float)rv[r];
                    }
                    bf16x4 q;
#pragma unroll
                    for (int r = 0; r < 4; ++r) q[r] = (bf16_t)v[r];
                    *reinterpret_cast<bf16x4*>(out + o + co0) = q;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (co0 + r >= Cout) break;
                        float x = v[r];
                        if (resid) x += (float)resid[o + co0 + r];
                        out[o + co0 + r] = (bf16_t)x;
                    }
                }
            }
        }
    }
}

// Persistent LDS variant for 64 input channels: the 64 -> 64 body of the network and its last
// layer (64 -> 1).  One 512-thread workgroup per CU keeps all nine taps of the weights in LDS
// (72 KB for 64 output channels) and walks 16 x 32 pixel
// tiles: the (16+2) x (32+2) input halo of the current tile sits in LDS (76.5 KB) while the
// next tile's halo is already loaded into registers.  Wave w computes rows 2w, 2w+1 of the
// tile: 4 MFMA row tiles (16 pixels) x 4 column tiles (16 channels), 16 MFMAs per 8 LDS
// operand reads in each K step of 32.  Rows of 8 16-byte chunks are stored with the chunk
// index XOR-ed by (row >> 1) & 7, so the operand reads of 16 consecutive pixels (or output
// channels) hit 16 different bank groups without padding.
constexpr int kPTileH = 16, kPTileW = 32;
constexpr int kPHaloW = kPTileW + 2;
constexpr int kPHaloPix = (kPTileH + 2) * kPHaloW;
constexpr int kPHaloChunks = kPHaloPix * 8;
constexpr int kPThreads = 512;
constexpr int kPPer = (kPHaloChunks + kPThreads - 1) / kPThreads;

__device__ inline int swz(int row, int chunk) { return chunk ^ ((row >> 1) & 7); }

// NT = 4 column tiles: 64 output channels, staged epilogue with 16-byte stores (and optional
// residual); NT = 1: up to 16 output channels (weights past Cout are zero), each lane stores its
// own channels (the network's last layer, 64 -> 1).
template <bool RELU, bool RESID, int NT>
__global__ void __launch_bounds__(kPThreads) conv3x3_fwd_persistent_kernel(
    const bf16_t* __restrict__ in, const bf16_t* __restrict__ w, const float* __restrict__ bias,
    const bf16_t* __restrict__ resid, bf16_t* __restrict__ out, int H, int W, int Cout, int tiles_x,
    int tiles_y, int n_tiles) {
    static_assert(NT == 4 || (NT == 1 && !RESID), "64 channels, or at most 16 without residual");
    constexpr int K = 9 * 64;
    constexpr int kCo = 16 * NT;  // output channels held in LDS (zero past Cout)
    __shared__ __attribute__((aligned(16))) bf16_t wl[kCo * 9 * 64];
    __shared__ __attribute__((aligned(16))) bf16_t halo[kPHaloPix * 64];
    const int tid = threadIdx.x;
    // all weights: row (co, tap) of 8 chunks
    for (int e = tid; e < kCo * 9 * 8; e += kPThreads) {
        const int c = e & 7, row = e >> 3;  // row = co * 9 + tap
        const int co = row / 9;
        *reinterpret_cast<bf16x8*>(&wl[(row * 8 + swz(co, c)) * 8]) =
            co < Cout ? *reinterpret_cast<const bf16x8*>(w + (size_t)co * K + (row - co * 9) * 64 + c * 8) : zero8();
    }
    // the halo chunks of this thread, the same for every tile: position in the halo relative to
    // the tile origin (hy, hx >= -1), element offset in the input, element offset in LDS
    int h_yx[kPPer], h_in[kPPer], h_lds[kPPer];
#pragma unroll
    for (int p = 0; p < kPPer; ++p) {
        const int e = tid + p * kPThreads;
        const int c = e & 7, pix = min(e >> 3, kPHaloPix - 1);
        const int hy = pix / kPHaloW - 1, hx = pix % kPHaloW - 1;
        h_yx[p] = ((hy + 1) << 16) | (hx + 1);
        h_in[p] = (hy * W + hx) * 64 + c * 8;
        h_lds[p] = (pix * 8 + swz(pix, c)) * 8;
    }
    bf16x8 pre[kPPer];
    auto fetch = [&](int tile) {
        const int tx = tile % tiles_x;
        const int ty = (tile / tiles_x) % tiles_y;
        const int n = tile / (tiles_x * tiles_y);
        const int ty0 = ty * kPTileH, tx0 = tx * kPTileW;
        const bf16_t* base = in + ((size_t)(n * H + ty0) * W + tx0) * 64;
#pragma unroll
        for (int p = 0; p < kPPer; ++p) {
            const int e = tid + p * kPThreads;
            const int yy = ty0 - 1 + (h_yx[p] >> 16), xx = tx0 - 1 + (h_yx[p] & 0xffff);
            bf16x8 v = zero8();
            if (e < kPHaloChunks && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W)
                v = *reinterpret_cast<const bf16x8*>(base + h_in[p]);
            pre[p] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < kPPer; ++p) {
            const int e = tid + p * kPThreads;
            if (e < kPHaloChunks) *reinterpret_cast<bf16x8*>(&halo[h_lds[p]]) = pre[p];
        }
    };
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    // the MFMAs take the weights as the row operand, so lane (m, g) holds output channels
    // 16 t + 4 g + r of pixel m: the bias of those channels starts every accumulator
    float bias_t[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = t * 16 + 4 * g + r;
            bias_t[t][r] = (bias && co < Cout) ? bias[co] : 0.0f;
        }
    int tile = blockIdx.x;
    if (tile < n_tiles) {
        fetch(tile);
        store();
    }
    __syncthreads();
    while (tile < n_tiles) {
        const int next = tile + (int)gridDim.x;
        f32x4 acc[4][NT];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[i][t] = f32x4{bias_t[t][0], bias_t[t][1], bias_t[t][2], bias_t[t][3]};
        // the previous tile's stores have completed (their data registers are reused by the
        // accumulators); the next tile's halo is then in flight while this tile is computed
        __builtin_amdgcn_s_waitcnt(0);
        if (next < n_tiles) fetch(next);
#pragma unroll 1
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int c0 = 0; c0 < 8; c0 += 4) {  // chunk offset of the K step (32 channels)
                const int ch = c0 + g;
                bf16x8 b[NT], a[4];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int co = t * 16 + m;
                    b[t] = *reinterpret_cast<const bf16x8*>(&wl[((co * 9 + tap) * 8 + swz(co, ch)) * 8]);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int pix = (2 * wave + (i >> 1) + dy) * kPHaloW + 16 * (i & 1) + m + dx;
                    a[i] = *reinterpret_cast<const bf16x8*>(&halo[(pix * 8 + swz(pix, ch)) * 8]);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[t], a[i], acc[i][t], 0, 0, 0);
            }
        }
        if constexpr (NT == 1) {
            // lane (m, g) holds channels 4 g .. 4 g + 3 of pixel m of each row tile: stored
            // straight from the accumulators (at most 16 channels per pixel)
            const int tx = tile % tiles_x;
            const int ty = (tile / tiles_x) % tiles_y;
            const int n = tile / (tiles_x * tiles_y);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int yy = ty * kPTileH + 2 * wave + (i >> 1), xx = tx * kPTileW + 16 * (i & 1) + m;
                if (yy >= H || xx >= W) continue;
                const size_t o = ((size_t)(n * H + yy) * W + xx) * Cout;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = 4 * g + r;
                    if (co < Cout) out[o + co] = (bf16_t)(RELU ? fmaxf(acc[i][0][r], 0.0f) : acc[i][0][r]);
                }
            }
        } else {
            const int tx = tile % tiles_x;
            const int ty = (tile / tiles_x) % tiles_y;
            const int n = tile / (tiles_x * tiles_y);
            // this thread's output chunks (pixel, 8 channels): their residual is loaded now, so
            // that it arrives while the results are staged
            constexpr int kChunks = kPTileH * kPTileW * 8 / kPThreads;
            bf16x8 rv[RESID ? kChunks : 1];
            size_t o[kChunks];
            bool ok[kChunks];
#pragma unroll
            for (int k = 0; k < kChunks; ++k) {
                const int e = tid + k * kPThreads;
                const int c = e & 7, p = e >> 3;
                const int yy = ty * kPTileH + p / kPTileW, xx = tx * kPTileW + p % kPTileW;
                ok[k] = yy < H && xx < W;
                o[k] = ((size_t)(n * H + min(yy, H - 1)) * W + min(xx, W - 1)) * 64 + c * 8;
                if constexpr (RESID) rv[k] = *reinterpret_cast<const bf16x8*>(resid + o[k]);
            }
            __syncthreads();  // every wave is done with this halo: reuse it to stage the results
            // (ReLU), as bfloat16, into stage[pixel][channel] (pixel = row * 32 + column, rows of 8
            // 16-byte chunks swizzled like the halo): four channels per 8-byte store
            bf16_t* stage = halo;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = (2 * wave + (i >> 1)) * kPTileW + 16 * (i & 1) + m;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int co = t * 16 + 4 * g;
                    bf16x4 q;
#pragma unroll
                    for (int r = 0; r < 4; ++r) q[r] = (bf16_t)(RELU ? fmaxf(acc[i][t][r], 0.0f) : acc[i][t][r]);
                    *reinterpret_cast<bf16x4*>(&stage[(p * 8 + swz(p, co >> 3)) * 8 + (co & 7)]) = q;
                }
            }
            __syncthreads();
            // 16-byte chunks (pixel, 8 channels); residual added in FP32
#pragma unroll
            for (int k = 0; k < kChunks; ++k) {
                const int e = tid + k * kPThreads;
                const int c = e & 7, p = e >> 3;
                bf16x8 v = *reinterpret_cast<const bf16x8*>(&stage[(p * 8 + swz(p, c)) * 8]);
                if constexpr (RESID) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) v[q] = (bf16_t)((float)v[q] + (float)rv[k][q]);
                }
                if (ok[k]) *reinterpret_cast<bf16x8*>(out + o[k]) = v;
            }
        }
        __syncthreads();  // every wave is done with this halo
        if (next < n_tiles) store();
        __syncthreads();
        tile = next;
    }
}

constexpr int kWgTile = 64;     // pixels per LDS tile: two MFMA K steps of 32
constexpr int kWgRow = 72;      // bf16 per LDS row (64 pixels + 8 of padding: conflict-free reads)

// Weight and bias gradients of one tap on MFMA: grad_w[co][ci] = sum_p gout[p][co] x[p + tap][ci],
// a 64 x 64 GEMM with K = pixels.  The workgroup of (pixel chunk, tap) stages 64-pixel tiles of
// the output gradient and of the tap-shifted input in LDS transposed to [channel][pixel] (bf16,
// rows padded to 72 so that the 16 rows a 16-lane group reads fall in distinct banks); wave w
// accumulates rows 16 w .. 16 w + 15 of grad_w against all 64 columns with
// v_mfma_f32_16x16x32_bf16 (lane l gives row/column l % 16 with pixels 8 (l / 16) .. +7 of the
// K step and receives rows 4 (l / 16) .. +3 of column l % 16), FP32 accumulation.  Partial
// sums per chunk are reduced in a fixed order (deterministic).
__global__ void __launch_bounds__(256) conv3x3_wgrad_kernel(
    const bf16_t* __restrict__ in, const bf16_t* __restrict__ gout, float* __restrict__ part_w,
    float* __restrict__ part_b, int N, int H, int W, int Cin, int Cout, int chunk_px) {
    __shared__ __attribute__((aligned(16))) bf16_t ty[kMaxCout * kWgRow];  // gout^T  [co][p]
    __shared__ __attribute__((aligned(16))) bf16_t tx[kMaxCout * kWgRow];  // shifted x^T [ci][p]
    const int tap = blockIdx.y;
    const int chunk = blockIdx.x;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int P = N * H * W;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float accb = 0.0f;
    const int p_begin = chunk * chunk_px;
    const int p_end = min(p_begin + chunk_px, P);
    // staging: thread tid loads channels c8 .. c8+7 of the pixel pair (2 q, 2 q + 1) of the tile
    const int q = tid >> 3, c8 = (tid & 7) * 8;
    const bool vec_y = (Cout & 7) == 0, vec_x = (Cin & 7) == 0;
    auto load8 = [&](const bf16_t* base, int C, int c0, bool vec) {
        bf16x8 v = zero8();
        if (vec) {
            if (c0 < C) v = *reinterpret_cast<const bf16x8*>(base + c0);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (c0 + e < C) v[e] = base[c0 + e];
        }
        return v;
    };
    for (int p0 = p_begin; p0 < p_end; p0 += kWgTile) {
        bf16x8 vy[2], vx[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p = p0 + 2 * q + h;
            vy[h] = zero8();
            vx[h] = zero8();
            if (p < p_end) {
                vy[h] = load8(gout + (size_t)p * Cout, Cout, c8, vec_y);
                const int x = p % W;
                const int t = p / W;
                const int y = t % H, n = t / H;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    vx[h] = load8(in + ((size_t)(n * H + yy) * W + xx) * Cin, Cin, c8, vec_x);
            }
        }
        // transposed store: channel c8 + e, pixels 2 q and 2 q + 1 in one 32-bit word
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int row = (c8 + e) * kWgRow + 2 * q;
            bf16_t py2[2] = {vy[0][e], vy[1][e]};
            bf16_t px2[2] = {vx[0][e], vx[1][e]};
            *reinterpret_cast<uint32_t*>(&ty[row]) = *reinterpret_cast<const uint32_t*>(py2);
            *reinterpret_cast<uint32_t*>(&tx[row]) = *reinterpret_cast<const uint32_t*>(px2);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kWgTile / 32; ++ks) {
            const int pk = ks * 32 + 8 * g;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&ty[(16 * wave + m) * kWgRow + pk]);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8 bb = *reinterpret_cast<const bf16x8*>(&tx[(16 * t + m) * kWgRow + pk]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
            }
        }
        if (tap == 0 && tid < kMaxCout) {
#pragma unroll 8
            for (int p = 0; p < kWgTile; ++p) accb += (float)ty[tid * kWgRow + p];
        }
        __syncthreads();
    }
    float* pw = part_w + ((size_t)chunk * 9 + tap) * kMaxCout * kMaxCout;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) pw[(16 * wave + 4 * g + r) * kMaxCout + 16 * t + m] = acc[t][r];
    if (tap == 0 && tid < kMaxCout) part_b[(size_t)chunk * kMaxCout + tid] = accb;
}

// grad_w[co][tap][ci] = sum over chunks (fixed order); grad_b[co] likewise
__global__ void conv3x3_wgrad_reduce(const float* __restrict__ part_w, const float* __restrict__ part_b,
                                     float* __restrict__ grad_w, float* __restrict__ grad_b,
                                     int n_chunks, int Cin, int Cout) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = Cout * 9 * Cin;
    if (idx < total) {
        const int ci = idx % Cin;
        const int tap = (idx / Cin) % 9;
        const int co = idx / (9 * Cin);
        float s = 0.0f;
        for (int c = 0; c < n_chunks; ++c)
            s += part_w[(((size_t)c * 9 + tap) * kMaxCout + co) * kMaxCout + ci];
        grad_w[idx] = s;
    }
    if (grad_b && idx < Cout) {
        float s = 0.0f;
        for (int c = 0; c < n_chunks; ++c) s += part_b[(size_t)c * kMaxCout + idx];
        grad_b[idx] = s;
    }
}

int conv_chunks(long long P) {
    long long c = (P + 2047) / 2048;
    if (c > 512) c = 512;
    if (c < 1) c = 1;
    return (int)c;
}

}  // namespace

extern "C" {

int ctr_conv3x3_bf16_forward(const void* in, const void* weight, const float* bias,
                             const void* residual, void* out, int n, int h, int w, int cin, int cout,
                             int relu, void* stream) {
    if (!in || !weight || !out) return CTR_ERR_INVALID;
    if (n <= 0 || h <= 0 || w <= 0 || cin <= 0 || cout <= 0 || cout > kMaxCout) return CTR_ERR_INVALID;
    const long long P = (long long)n * h * w;
    if (P * (cin > cout ? cin : cout) >= 2147483647LL) return CTR_ERR_INVALID;
    const long long blocks = (P + kConvWaves * kPxPerWave - 1) / (kConvWaves * kPxPerWave);
    hipStream_t s = (hipStream_t)stream;
    const bf16_t* x = (const bf16_t*)in;
    const bf16_t* wt = (const bf16_t*)weight;
    const bf16_t* r = (const bf16_t*)residual;
    bf16_t* y = (bf16_t*)out;
    // the generic kernel stores four channels at a time when rows allow it
    const int vec = (cout % 4 == 0) && ((reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(residual)) & 7) == 0;
    if (cin == 64 && (cout == 64 || (cout <= 16 && !r))) {
        const int tiles_x = (w + kPTileW - 1) / kPTileW, tiles_y = (h + kPTileH - 1) / kPTileH;
        const long long n_tiles = (long long)n * tiles_x * tiles_y;
        if (n_tiles >= 2147483647LL) return CTR_ERR_INVALID;
        int dev = 0, n_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return CTR_ERR_HIP;
        const int grid = (int)std::min<long long>(n_tiles, std::max(n_cu, 1));
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kPThreads), 0, s, x, wt, bias, r, y, h, w, cout, tiles_x,
                               tiles_y, (int)n_tiles);
        };
        if (cout <= 16)
            relu ? launch(conv3x3_fwd_persistent_kernel<true, false, 1>) : launch(conv3x3_fwd_persistent_kernel<false, false, 1>);
        else if (relu)
            r ? launch(conv3x3_fwd_persistent_kernel<true, true, 4>) : launch(conv3x3_fwd_persistent_kernel<true, false, 4>);
        else
            r ? launch(conv3x3_fwd_persistent_kernel<false, true, 4>) : launch(conv3x3_fwd_persistent_kernel<false, false, 4>);
    } else if (cin <= 4) {
        const int tiles_x = (w + kSTileW - 1) / kSTileW, tiles_y = (h + kSTileH - 1) / kSTileH;
        const long long n_tiles = (long long)n * tiles_x * tiles_y;
        if (n_tiles >= 2147483647LL) return CTR_ERR_INVALID;
        int dev = 0, n_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return CTR_ERR_HIP;
        // persistent: eight workgroups per CU walk the tiles
        const int grid = (int)std::min<long long>(n_tiles, 8LL * std::max(n_cu, 1));
        hipLaunchKernelGGL(conv3x3_fwd_small_cin_kernel, dim3(grid), dim3(256), 0, s, x, wt, bias, r, y,
                           h, w, cin, cout, relu, vec, tiles_x, tiles_y, (int)n_tiles);
    } else if (cin % 32 == 0) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(conv3x3_fwd_kernel<true>), dim3((unsigned)blocks), dim3(256),
                           0, s, x, wt, bias, r, y, n, h, w, cin, cout, relu, vec);
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(conv3x3_fwd_kernel<false>), dim3((unsigned)blocks), dim3(256),
                           0, s, x, wt, bias, r, y, n, h, w, cin, cout, relu, vec);
    }
    return hipGetLastError() == hipSuccess ? CTR_OK : CTR_ERR_HIP;
}

size_t ctr_conv3x3_bf16_wgrad_workspace_bytes(int n, int h, int w, int cin, int cout) {
    (void)cin;
    (void)cout;
    if (n <= 0 || h <= 0 || w <= 0) return 0;
    const int chunks = conv_chunks((long long)n * h * w);
    return (size_t)chunks * (9 * kMaxCout * kMaxCout + kMaxCout) * sizeof(float);
}

int ctr_conv3x3_bf16_wgrad(const void* in, const void* grad_out, float* grad_weight, float* grad_bias,
                           int n, int h, int w, int cin, int cout, void* workspace,
                           size_t workspace_bytes, void* stream) {
    if (!in || !grad_out || !grad_weight || !workspace) return CTR_ERR_INVALID;
    if (n <= 0 || h <= 0 || w <= 0 || cin <= 0 || cout <= 0 || cin > kMaxCout || cout > kMaxCout)
        return CTR_ERR_INVALID;
    const long long P = (long long)n * h * w;
    if (P * kMaxCout >= 2147483647LL) return CTR_ERR_INVALID;
    if (workspace_bytes < ctr_conv3x3_bf16_wgrad_workspace_bytes(n, h, w, cin, cout)) return CTR_ERR_WORKSPACE;
    const int chunks = conv_chunks(P);
    const int chunk_px = (int)(((P + chunks - 1) / chunks + kWgTile - 1) / kWgTile * kWgTile);
    float* part_w = (float*)workspace;
    float* part_b = part_w + (size_t)chunks * 9 * kMaxCout * kMaxCout;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(conv3x3_wgrad_kernel, dim3(chunks, 9), dim3(256), 0, s, (const bf16_t*)in,
                       (const bf16_t*)grad_out, part_w, part_b, n, h, w, cin, cout, chunk_px);
    if (hipGetLastError() != hipSuccess) return CTR_ERR_HIP;
    const int total = cout * 9 * cin;
    hipLaunchKernelGGL(conv3x3_wgrad_reduce, dim3((total + 255) / 256), dim3(256), 0, s, part_w, part_b,
                       grad_weight, grad_bias, chunks, cin, cout);
    return hipGetLastError() == hipSuccess ? CTR_OK : CTR_ERR_HIP;
}

}  // extern "C"
