// Learned psi extrapolation (a chosen form with no counterpart in the reference): a 3x3
// convolutional network whose layers all have 32 channels (the 3 input and 2 output channels are
// zero-padded to 32), BF16 weights and activations in NHWC layout, FP32 accumulation on the matrix
// cores (v_mfma_f32_32x32x16_bf16), periodic padding.  Forward, backward-data and backward-weight
// kernels, the optimiser and the glue between the psi history and the network.
//
// MFMA 32x32x16 operands (wave of 64 lanes, i = lane % 32, h = lane / 32): lane (i, h) gives row
// m = i of A and column n = i of B, eight bf16 each (operand slots e = 0..7); slot e of lane (i, h)
// in A and in B meet in the same K index, so the kernels put the same pixel or channel, 8 h + e of
// the K-block, into both and the product sums over the block whatever the instruction's K order.
// D[m][n] with n = i and m = 8 (r / 4) + 4 h + r % 4 for the lane's 16 results r.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "cmm_internal.hpp"
#include "../../include/cmm.h"

namespace cmm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kCh = 32;  // channels of every layer

// round to nearest even; NaN stays NaN
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : v >= n ? v - n : v; }

// out[n][y][x][co] = act(b[co] + sum_{t, ci} W[co][t][ci] in[n][y + dy_t][x + dx_t][ci]), tap
// t = 3 (dy + 1) + (dx + 1), periodic in y and x.  A block takes kCvRows rows of up to SEGS
// 32-pixel segments: its input rows y0 - 1 .. y0 + kCvRows (with the periodic halo pixels) go
// into LDS once, with coalesced 16-byte loads, at a pixel pitch of 80 bytes (the 16 lanes of a
// 16-byte LDS read pass then cover all 64 banks once).  Wave w computes row y0 + w: per segment
// M = output channels, N = pixels, K = 9 taps x 32 channels (18 MFMAs, B from LDS), the bias as
// the accumulators' start; its A operands (18 weight fragments) stay in registers.  A lane's 16
// results are pixel x0 + i, output channels 8 g + 4 h .. 8 g + 4 h + 3 for g = 0..3: four runs of
// 4 consecutive channels, written with 8-byte (16-byte for FP32) stores.  MODE 0: bf16 out with
// ReLU, 1: bf16 out, 2: fp32 out, 3: bf16 out, zero where mask (bf16, same layout as out) is not
// positive.  The backward-data pass is this kernel on the output gradient with the weights
// transposed in channels and flipped in taps; mode 3 also applies the ReLU derivative of the layer
// below.
constexpr int kCvRows = 4;    // output rows per block (one per wave)
constexpr int kCvPitch = 40;  // LDS elements per staged pixel (32 channels + 8 pad)

// SEGS: 32-pixel segments per block and row (launched with 1: the images too small for
// k_conv3_ring's strips of 4, where blocks of 4 segments would leave the device idle)
template <int MODE, int SEGS>
__global__ void __launch_bounds__(256) k_conv3(const unsigned short *__restrict__ in, const unsigned short *__restrict__ W,
                                               const float *__restrict__ bias, void *__restrict__ out, int H, int Wd,
                                               int groups, const unsigned short *__restrict__ mask) {
    constexpr int kCvPix = 32 * SEGS + 2;  // staged pixels per row
    __shared__ __align__(16) unsigned short tile[(kCvRows + 2) * kCvPix * kCvPitch];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int y0 = (blockIdx.x / groups) * kCvRows;
    const int x0 = (blockIdx.x % groups) * (32 * SEGS);
    const int nseg = min(SEGS, (Wd - x0) / 32);
    const int npix = 32 * nseg + 2;
    const long img = (long)blockIdx.y * H * Wd;
    // stage rows y0 - 1 .. y0 + kCvRows, pixels x0 - 1 .. x0 + 32 nseg: chunk c is 8 channels
    const int nchunk = (kCvRows + 2) * npix * 4;
    for (int c = threadIdx.x; c < nchunk; c += 256) {
        const int q = c & 3, px = (c >> 2) % npix, r = (c >> 2) / npix;
        const int yy = ((y0 - 1 + r) % H + H) % H;
        const int xx = wrapi(x0 - 1 + px, Wd);
        const uint4 v = *reinterpret_cast<const uint4 *>(in + (img + (long)yy * Wd + xx) * kCh + 8 * q);
        *reinterpret_cast<uint4 *>(tile + (r * kCvPix + px) * kCvPitch + 8 * q) = v;
    }
    const int i = lane & 31, h = lane >> 5;
    // A = weights (row m = output channel i), B = input (column n = pixel x0 + i); operand slot e of
    // lane (i, h) in K-block kb of tap t: input channel 16 kb + 8 h + e in both
    bf16x8 wf[18];
#pragma unroll
    for (int q = 0; q < 18; ++q)
        wf[q] = *reinterpret_cast<const bf16x8 *>(W + ((long)i * 9 + q / 2) * kCh + 8 * h + 16 * (q % 2));
    __syncthreads();
    const int y = y0 + wid;
    if (y >= H) return;  // uniform over the wave; no barrier follows
    for (int sg = 0; sg < nseg; ++sg) {
        const long o = (img + (long)y * Wd + x0 + 32 * sg + i) * kCh + 4 * h;  // + 8 g: channel run g
        f32x16 acc;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 bv = bias ? *reinterpret_cast<const float4 *>(bias + 8 * g + 4 * h) : float4{0.f, 0.f, 0.f, 0.f};
            acc[4 * g] = bv.x;
            acc[4 * g + 1] = bv.y;
            acc[4 * g + 2] = bv.z;
            acc[4 * g + 3] = bv.w;
        }
        uint2 mk[4];
        if (MODE == 3)
#pragma unroll
            for (int g = 0; g < 4; ++g) mk[g] = *reinterpret_cast<const uint2 *>(mask + o + 8 * g);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            // staged row wid + 1 + dy, staged pixel 32 sg + i + dx + 1
            const unsigned short *bp = tile + ((wid + t / 3) * kCvPix + 32 * sg + i + t % 3) * kCvPitch + 8 * h;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const bf16x8 bx = *reinterpret_cast<const bf16x8 *>(bp + 16 * kb);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[2 * t + kb], bx, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = acc[4 * g + j];
                if (MODE == 0) v[j] = fmaxf(v[j], 0.f);
                if (MODE == 3) {
                    const unsigned a = ((j < 2 ? mk[g].x : mk[g].y) >> (16 * (j & 1))) & 0xffffu;
                    if ((a & 0x8000u) || (a & 0x7fffu) == 0) v[j] = 0.f;
                }
            }
            if (MODE == 2) {
                *reinterpret_cast<float4 *>(static_cast<float *>(out) + o + 8 * g) = float4{v[0], v[1], v[2], v[3]};
            } else {
                const uint2 pk = {(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16),
                                  (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
                *reinterpret_cast<uint2 *>(static_cast<unsigned short *>(out) + o + 8 * g) = pk;
            }
        }
    }
}

// k_conv3 for large images with blocks that stay over a strip of 4 segments (128 pixels) and a
// range of row groups: the input rows of the group being computed sit in a ring of 6 LDS rows
// (row r in slot (r + 1) mod 6), and while the group's products run, the next group's 4 new rows
// come from global memory into registers; after a barrier they replace the 4 rows no longer
// needed.  Same products and epilogue as k_conv3 (bit for bit).
template <int MODE>
__global__ void __launch_bounds__(256) k_conv3_ring(const unsigned short *__restrict__ in,
                                                    const unsigned short *__restrict__ W, const float *__restrict__ bias,
                                                    void *__restrict__ out, int H, int Wd, int strips, int rgroups,
                                                    int gper, const unsigned short *__restrict__ mask) {
    constexpr int SEGS = 4, PIX = 32 * SEGS + 2;
    constexpr int NPF = (4 * PIX * 4 + 255) / 256;  // 16-byte chunks of 4 rows per thread
    __shared__ __align__(16) unsigned short tile[6 * PIX * kCvPitch];
    const int g0 = (blockIdx.x / strips) * gper, g1 = min(rgroups, g0 + gper);
    if (g0 >= g1) return;  // uniform over the block, before any barrier
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int x0 = (blockIdx.x % strips) * (32 * SEGS);
    const int nseg = min(SEGS, (Wd - x0) / 32);
    const int npix = 32 * nseg + 2;
    const long img = (long)blockIdx.y * H * Wd;
    auto slot = [](int r) { return (r + 7) % 6; };  // (r + 1) mod 6 for r >= -1
    auto src = [=](int r, int px, int q) {
        const int yy = (r % H + H) % H, xx = wrapi(x0 - 1 + px, Wd);
        return in + (img + (long)yy * Wd + xx) * kCh + 8 * q;
    };
    // the first group's 6 rows
    {
        const int ya = 4 * g0;
        for (int c = threadIdx.x; c < 6 * npix * 4; c += 256) {
            const int q = c & 3, px = (c >> 2) % npix, r = ya - 1 + (c >> 2) / npix;
            *reinterpret_cast<uint4 *>(tile + (slot(r) * PIX + px) * kCvPitch + 8 * q) =
                *reinterpret_cast<const uint4 *>(src(r, px, q));
        }
    }
    const int i = lane & 31, h = lane >> 5;
    bf16x8 wf[18];
#pragma unroll
    for (int q = 0; q < 18; ++q)
        wf[q] = *reinterpret_cast<const bf16x8 *>(W + ((long)i * 9 + q / 2) * kCh + 8 * h + 16 * (q % 2));
    __syncthreads();
    for (int g = g0; g < g1; ++g) {
        const int ya = 4 * g;
        const bool more = g + 1 < g1;  // uniform over the block
        u32x4 pf[NPF];  // a native vector: struct copies would go through private memory
#pragma unroll
        for (int j = 0; j < NPF; ++j) {
            // a chunk past the 4 rows repeats the last one (after the last group: chunk 0, unused)
            const int c = min(threadIdx.x + 256 * j, 4 * npix * 4 - 1) * (more ? 1 : 0);
            const int q = c & 3, px = (c >> 2) % npix, r = ya + 5 + (c >> 2) / npix;
            pf[j] = *reinterpret_cast<const u32x4 *>(src(r, px, q));
        }
        const int y = ya + wid;
        if (y < H) {
            for (int sg = 0; sg < nseg; ++sg) {
                const long o = (img + (long)y * Wd + x0 + 32 * sg + i) * kCh + 4 * h;
                f32x16 acc;
#pragma unroll
                for (int gg = 0; gg < 4; ++gg) {
                    const float4 bv =
                        bias ? *reinterpret_cast<const float4 *>(bias + 8 * gg + 4 * h) : float4{0.f, 0.f, 0.f, 0.f};
                    acc[4 * gg] = bv.x;
                    acc[4 * gg + 1] = bv.y;
                    acc[4 * gg + 2] = bv.z;
                    acc[4 * gg + 3] = bv.w;
                }
                uint2 mk[4];
                if (MODE == 3)
#pragma unroll
                    for (int gg = 0; gg < 4; ++gg) mk[gg] = *reinterpret_cast<const uint2 *>(mask + o + 8 * gg);
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const unsigned short *bp =
                        tile + (slot(y + t / 3 - 1) * PIX + 32 * sg + i + t % 3) * kCvPitch + 8 * h;
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb) {
                        const bf16x8 bx = *reinterpret_cast<const bf16x8 *>(bp + 16 * kb);
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[2 * t + kb], bx, acc, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int gg = 0; gg < 4; ++gg) {
                    float v[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v[j] = acc[4 * gg + j];
                        if (MODE == 0) v[j] = fmaxf(v[j], 0.f);
                        if (MODE == 3) {
                            const unsigned a = ((j < 2 ? mk[gg].x : mk[gg].y) >> (16 * (j & 1))) & 0xffffu;
                            if ((a & 0x8000u) || (a & 0x7fffu) == 0) v[j] = 0.f;
                        }
                    }
                    if (MODE == 2) {
                        *reinterpret_cast<float4 *>(static_cast<float *>(out) + o + 8 * gg) = float4{v[0], v[1], v[2], v[3]};
                    } else {
                        const uint2 pk = {(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16),
                                          (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
                        *reinterpret_cast<uint2 *>(static_cast<unsigned short *>(out) + o + 8 * gg) = pk;
                    }
                }
            }
        }
        __syncthreads();  // the group's reads of the ring are done
        if (more) {
            // the same chunk indices as the loads: a chunk past the 4 rows repeats the last one,
            // the same data to the same place
#pragma unroll
            for (int j = 0; j < NPF; ++j) {
                const int c = min(threadIdx.x + 256 * j, 4 * npix * 4 - 1);
                const int q = c & 3, px = (c >> 2) % npix, r = ya + 5 + (c >> 2) / npix;
                *reinterpret_cast<u32x4 *>(tile + (slot(r) * PIX + px) * kCvPitch + 8 * q) = pf[j];
            }
            __syncthreads();
        }
    }
}

// dW[co][t][ci] += sum_p g[p][co] in[p + off_t][ci], db[co] += sum_p g[p][co].  A unit of work
// is a 32-pixel column segment of kWgRows rows of g in one image; a wave takes units in turn
// (stride: all waves of the grid) and keeps the 9 taps' 32 x 32 sums in its accumulators.  It
// walks the input rows r = y0 - 1 .. y1 of its unit: row r's 34 pixels (with the periodic halo)
// and g row r + 1 go from global memory (16-byte loads, the next row's issued before this row's
// products) into the wave's own LDS tiles, transposed to [channel][pixel] so that a lane reads the
// 8 consecutive pixels of one channel that an MFMA operand takes; the input row is written three
// times, shifted by dx = -1, 0, 1.  g rows r - 1, r, r + 1 stay in a window of three tiles, so
// every g row and input row is read from global memory once per unit.  Per row: 2 K-blocks of 16
// pixels x (3 taps dy) x (3 taps dx) v_mfma_f32_32x32x16_bf16 (M = output channels from g, N =
// input channels, K = pixels; A and B take the same pixel for the same operand slot, so the sum is
// over pixels whatever the instruction's K order).  The block's 4 waves add their sums through LDS
// and the block adds them to dW and db with FP32 atomics.
constexpr int kWgRows = 32;              // g rows per unit
constexpr int kWgPitch = 40;             // LDS pitch (pixels) of a transposed tile
constexpr int kWgTile = kCh * kWgPitch;  // elements of one tile
constexpr int kWgWaves = 4;

__device__ __forceinline__ void wg_tile_put(unsigned short *t, int p, int q, const uint4 v) {
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) t[(8 * q + j) * kWgPitch + p] = (unsigned short)(w[j >> 1] >> (16 * (j & 1)));
}

__global__ void __launch_bounds__(64 * kWgWaves) __attribute__((amdgpu_waves_per_eu(2))) k_conv3_wgrad(const unsigned short *__restrict__ g,
                                                               const unsigned short *__restrict__ in,
                                                               float *__restrict__ dW, float *__restrict__ db, int H,
                                                               int Wd, int bands, long units) {
    __shared__ __align__(16) unsigned short lds[kWgWaves * 6 * kWgTile];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned short *gT = lds + wid * 6 * kWgTile;  // 3 g rows
    unsigned short *iT = gT + 3 * kWgTile;         // the input row, shifted by dx = -1, 0, 1
    const int i = lane & 31, h = lane >> 5;
    const int segs = Wd / 32;
    // this lane's 16-byte chunks of a 32-pixel row: pixels p0 and p0 + 16, channels 8q .. 8q + 7;
    // lanes 0..7 also take the halo pixels (x0 - 1 for lanes 0..3, x0 + 32 for lanes 4..7)
    const int p0 = lane >> 2, q = lane & 3;
    f32x16 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x16{};
    float bs[8] = {};
    for (long u = (long)blockIdx.x * kWgWaves + wid; u < units; u += (long)gridDim.x * kWgWaves) {
        const int seg = (int)(u % segs);
        const long rest = u / segs;
        const int band = (int)(rest % bands);
        const long n = rest / bands;
        const int x0 = 32 * seg, y0 = band * kWgRows, y1 = min(H, y0 + kWgRows);
        const int xh = lane < 4 ? wrapi(x0 - 1, Wd) : wrapi(x0 + 32, Wd);
        const unsigned short *gimg = g + n * H * Wd * kCh;
        const unsigned short *iimg = in + n * H * Wd * kCh;
        uint4 gv[2], iv[2], hv = {};
        auto load = [&](int r) {
            const unsigned short *irow = iimg + (long)wrapi(r, H) * Wd * kCh;
#pragma unroll
            for (int k = 0; k < 2; ++k)
                iv[k] = *reinterpret_cast<const uint4 *>(irow + (long)(x0 + p0 + 16 * k) * kCh + 8 * q);
            if (lane < 8) hv = *reinterpret_cast<const uint4 *>(irow + (long)xh * kCh + 8 * q);
            if (r + 1 < y1) {
                const unsigned short *grow = gimg + (long)(r + 1) * Wd * kCh;
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    gv[k] = *reinterpret_cast<const uint4 *>(grow + (long)(x0 + p0 + 16 * k) * kCh + 8 * q);
            }
        };
        load(y0 - 1);
        for (int r = y0 - 1; r <= y1; ++r) {
            // this row into the tiles (the previous row's reads of them are done: LDS operations of
            // a wave complete in order, and the fence keeps the compiler from moving them)
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int p = p0 + 16 * k;
                if (p <= 30) wg_tile_put(iT, p + 1, q, iv[k]);  // dx = -1: position m holds pixel x0 + m - 1
                wg_tile_put(iT + kWgTile, p, q, iv[k]);
                if (p >= 1) wg_tile_put(iT + 2 * kWgTile, p - 1, q, iv[k]);
            }
            if (lane < 4) wg_tile_put(iT, 0, q, hv);
            else if (lane < 8) wg_tile_put(iT + 2 * kWgTile, 31, q, hv);
            if (r + 1 < y1) {
                unsigned short *gs = gT + ((r + 1 - y0) % 3) * kWgTile;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    wg_tile_put(gs, p0 + 16 * k, q, gv[k]);
                    const unsigned w[4] = {gv[k].x, gv[k].y, gv[k].z, gv[k].w};
#pragma unroll
                    for (int j = 0; j < 8; ++j) bs[j] += bf2f((unsigned short)(w[j >> 1] >> (16 * (j & 1))));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (r < y1) load(r + 1);  // the next row in flight during this row's products
#pragma unroll
            for (int kb = 0; kb < 32; kb += 16) {
                const int m0 = kb + 8 * h;
                bf16x8 B[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) B[c] = *reinterpret_cast<const bf16x8 *>(iT + c * kWgTile + i * kWgPitch + m0);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int y = r - (d - 1);  // g row of tap dy = d - 1
                    if (y < y0 || y >= y1) continue;
                    const bf16x8 A = *reinterpret_cast<const bf16x8 *>(gT + ((y - y0) % 3) * kWgTile + i * kWgPitch + m0);
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        acc[3 * d + c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B[c], acc[3 * d + c], 0, 0, 0);
                }
            }
        }
    }
    // bias sums: lanes with the same channel group q add up
#pragma unroll
    for (int j = 0; j < 8; ++j)
        for (int o = 4; o < 64; o <<= 1) bs[j] += __shfl_xor(bs[j], o, 64);
    __syncthreads();
    float *red = reinterpret_cast<float *>(lds);  // kWgWaves x 1024 sums, then kWgWaves x 32 bias sums
#pragma unroll
    for (int t = 0; t < 9; ++t) {
#pragma unroll
        for (int r = 0; r < 16; ++r) red[wid * 1024 + (8 * (r >> 2) + 4 * h + (r & 3)) * kCh + i] = acc[t][r];
        __syncthreads();
        for (int e = threadIdx.x; e < 1024; e += 64 * kWgWaves) {
            float v = red[e];
#pragma unroll
            for (int w = 1; w < kWgWaves; ++w) v += red[w * 1024 + e];
            atomicAdd(&dW[((long)(e >> 5) * 9 + t) * kCh + (e & 31)], v);
        }
        __syncthreads();
    }
    if (lane < 4)
#pragma unroll
        for (int j = 0; j < 8; ++j) red[wid * 32 + 8 * q + j] = bs[j];
    __syncthreads();
    if (threadIdx.x < kCh) {
        float v = red[threadIdx.x];
#pragma unroll
        for (int w = 1; w < kWgWaves; ++w) v += red[w * 32 + threadIdx.x];
        atomicAdd(&db[threadIdx.x], v);
    }
}

// g = g where the layer's output was positive (ReLU derivative)
__global__ void __launch_bounds__(256) k_relu_mask(long n, const unsigned short *__restrict__ act,
                                                   unsigned short *__restrict__ g) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const unsigned short a = act[j];
    if ((a & 0x8000u) || (a & 0x7fffu) == 0) g[j] = 0;
}

__global__ void __launch_bounds__(256) k_fill_f32(long n, float *__restrict__ p, float v) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j < n) p[j] = v;
}

// bf16 copies of the FP32 master weights of n_layers layers (each 32 x 9 x 32, W[co][t][ci]) and
// their transposed, tap-flipped copies Wt[ci][8 - t][co] for the backward-data pass
__global__ void __launch_bounds__(256) k_weights_bf16(long n, const float *__restrict__ w32,
                                                      unsigned short *__restrict__ w, unsigned short *__restrict__ wt) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const long per = (long)kCh * 9 * kCh;
    const long l = j / per;
    const int r = (int)(j - l * per);
    const int co = r / (9 * kCh), t = (r / kCh) % 9, ci = r % kCh;
    const unsigned short v = f2bf(w32[j]);
    w[j] = v;
    wt[l * per + ((long)ci * 9 + (8 - t)) * kCh + co] = v;
}

// AdamW (decoupled weight decay) on FP32 parameters; the gradient is scaled by gscale first
__global__ void __launch_bounds__(256) k_adamw(long n, float *__restrict__ p, const float *__restrict__ g,
                                               float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                               float b2, float eps, float wd, float gscale, float bc1, float bc2) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const float gj = g[j] * gscale;
    const float mj = b1 * m[j] + (1.f - b1) * gj;
    const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float mh = mj / bc1, vh = vj / bc2;
    p[j] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[j]);
}

// network input from three consecutive psi value planes (newest first): psi_n / s0,
// (psi_n - psi_{n-1}) / s1, (psi_n - 2 psi_{n-1} + psi_{n-2}) / s2 in channels 0..2, zeros after
__global__ void __launch_bounds__(256) k_cnn_input(long n, const double *__restrict__ p0, const double *__restrict__ p1,
                                                   const double *__restrict__ p2, const float *__restrict__ scales,
                                                   unsigned short *__restrict__ out) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double a = p0[j], b = p1[j], c = p2[j];
    const float s0 = scales[0], s1 = scales[1], s2 = scales[2];
    const unsigned c0 = f2bf((float)(a / s0)), c1 = f2bf((float)((a - b) / s1)),
                   c2 = f2bf((float)((a - 2.0 * b + c) / s2));
    // the pixel's 32 channels (64 bytes) as four 16-byte stores
    uint4 *o = reinterpret_cast<uint4 *>(out + j * kCh);
    o[0] = uint4{c0 | (c1 << 16), c2, 0u, 0u};
    o[1] = uint4{0u, 0u, 0u, 0u};
    o[2] = uint4{0u, 0u, 0u, 0u};
    o[3] = uint4{0u, 0u, 0u, 0u};
}

// Lagrange-weighted extrapolation of three psi value planes plus an optional correction
__global__ void __launch_bounds__(256) k_extrapolate(long n, const double *__restrict__ p0, const double *__restrict__ p1,
                                                     const double *__restrict__ p2, double w0, double w1, double w2,
                                                     const double *__restrict__ d, double *__restrict__ out) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double v = w0 * p0[j];
    v += w1 * p1[j];
    v += w2 * p2[j];
    if (d) v += d[j];
    out[j] = v;
}

// training targets: residuals of the Lagrange extrapolations to t_{n+1} and t_n + dt/2, scaled
__global__ void __launch_bounds__(256) k_cnn_target(long n, const double *__restrict__ p0, const double *__restrict__ p1,
                                                    const double *__restrict__ p2, const double *__restrict__ pn,
                                                    const double *__restrict__ ph, double a0, double a1, double a2,
                                                    double c0, double c1, double c2, const float *__restrict__ rs,
                                                    float *__restrict__ tgt) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const float r1 = rs[0], r12 = rs[1];
    double e1 = a0 * p0[j];
    e1 += a1 * p1[j];
    e1 += a2 * p2[j];
    double e12 = c0 * p0[j];
    e12 += c1 * p1[j];
    e12 += c2 * p2[j];
    tgt[2 * j] = (float)((pn[j] - e1) / r1);
    tgt[2 * j + 1] = (float)((ph[j] - e12) / r12);
}

// squared-error loss of the two predicted channels and its gradient (pred - target) in bf16;
// the sum of squares is accumulated per block and added once
__global__ void __launch_bounds__(256) k_cnn_mse(long n, const float *__restrict__ pred, const float *__restrict__ tgt,
                                                 unsigned short *__restrict__ g, double *__restrict__ loss) {
    // grid-stride: a capped grid adds one partial sum per block to *loss (one address: every
    // block's atomic add waits its turn)
    double s = 0;
    for (long j = (long)blockIdx.x * 256 + threadIdx.x; j < n; j += (long)gridDim.x * 256) {
        const float2 p = *reinterpret_cast<const float2 *>(pred + j * kCh);
        const float2 t = *reinterpret_cast<const float2 *>(tgt + 2 * j);
        const float d0 = p.x - t.x, d1 = p.y - t.y;
        s += (double)d0 * d0;
        s += (double)d1 * d1;
        // the pixel's 32 gradient channels (64 bytes) as four 16-byte stores
        uint4 *o = reinterpret_cast<uint4 *>(g + j * kCh);
        o[0] = uint4{(unsigned)f2bf(d0) | ((unsigned)f2bf(d1) << 16), 0u, 0u, 0u};
        o[1] = uint4{0u, 0u, 0u, 0u};
        o[2] = uint4{0u, 0u, 0u, 0u};
        o[3] = uint4{0u, 0u, 0u, 0u};
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

// correction value fields from the network output: d1 = r1 pred[.][0], d12 = r12 pred[.][1]
__global__ void __launch_bounds__(256) k_cnn_correction(long n, const float *__restrict__ pred,
                                                        const float *__restrict__ rs, double *__restrict__ d1,
                                                        double *__restrict__ d12) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const float r1 = rs[0], r12 = rs[1];
    d1[j] = (double)r1 * pred[j * kCh];
    d12[j] = (double)r12 * pred[j * kCh + 1];
}

// sum (a - b)^2 and sum b^2 (two doubles, accumulated)
__global__ void __launch_bounds__(256) k_dev_norms(long n, const double *__restrict__ a, const double *__restrict__ b,
                                                   double *__restrict__ out) {
    double s = 0, t = 0;
    for (long j = (long)blockIdx.x * 256 + threadIdx.x; j < n; j += (long)gridDim.x * 256) {
        const double d = a[j] - b[j];
        s += d * d;
        t += b[j] * b[j];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        t += __shfl_xor(t, off, 64);
    }
    __shared__ double red[8];
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = s;
        red[4 + (threadIdx.x >> 6)] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
        atomicAdd(out + 1, red[4] + red[5] + red[6] + red[7]);
    }
}

__global__ void k_zero2(double *p) {
    p[0] = 0.0;
    p[1] = 0.0;
}

static unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }
static int lc(const char *what) { return hip_check(hipGetLastError(), what); }

}  // namespace cmm

using namespace cmm;

namespace {
hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
bool img_ok(int n_img, int H, int W) { return n_img > 0 && H > 0 && W >= 32 && W % 32 == 0; }
}  // namespace

extern "C" {

namespace {
int conv3(int n_img, int H, int W, const uint16_t *in, const uint16_t *w, const float *b, int mode, void *out,
          const uint16_t *mask, void *stream) {
    // 4 segments per block where that still gives 1024 blocks or more
    const long rows = (H + kCvRows - 1) / kCvRows;
    const bool four = rows * ((W / 32 + 3) / 4) * n_img >= 1024;
    const int groups = four ? (W / 32 + 3) / 4 : W / 32;  // segment groups per row
    const dim3 grid((unsigned)(rows * groups), (unsigned)n_img);
#define CMM_CONV(M, SG) \
    hipLaunchKernelGGL((k_conv3<M, SG>), grid, dim3(256), 0, S(stream), in, w, b, out, H, W, groups, mask)
#define CMM_CONV_MODES(SG)               \
    if (mode == 0) CMM_CONV(0, SG);      \
    else if (mode == 1) CMM_CONV(1, SG); \
    else if (mode == 2) CMM_CONV(2, SG); \
    else CMM_CONV(3, SG);
    if (four) {
        // blocks over strips of 128 pixels and ranges of row groups, about 2 per CU
        const int strips = (W / 32 + 3) / 4;
        const int rgroups = (H + kCvRows - 1) / kCvRows;
        const long per = (long)strips * n_img;
        const int parts = (int)std::max<long>(1, std::min<long>(rgroups, 512 / per));
        const int gper = (rgroups + parts - 1) / parts;
        const dim3 gr((unsigned)(strips * parts), (unsigned)n_img);
#define CMM_RING(M) \
    hipLaunchKernelGGL((k_conv3_ring<M>), gr, dim3(256), 0, S(stream), in, w, b, out, H, W, strips, rgroups, gper, mask)
        if (mode == 0) CMM_RING(0);
        else if (mode == 1) CMM_RING(1);
        else if (mode == 2) CMM_RING(2);
        else CMM_RING(3);
#undef CMM_RING
    } else {
        CMM_CONV_MODES(1)
    }
#undef CMM_CONV_MODES
#undef CMM_CONV
    return lc("k_conv3");
}
}  // namespace

int cmm_cnn_forward(int n_img, int H, int W, const uint16_t *in, const uint16_t *w, const float *b, int mode, void *out,
                    void *stream) {
    if (!img_ok(n_img, H, W) || !in || !w || !out || mode < 0 || mode > 2) {
        set_error("cmm_cnn_forward: bad argument (width a multiple of 32)");
        return CMM_ERR_ARG;
    }
    return conv3(n_img, H, W, in, w, b, mode, out, nullptr, stream);
}

int cmm_cnn_backward_data(int n_img, int H, int W, const uint16_t *g, const uint16_t *wt, uint16_t *gin, void *stream) {
    return cmm_cnn_forward(n_img, H, W, g, wt, nullptr, 1, gin, stream);
}

int cmm_cnn_backward_data_relu(int n_img, int H, int W, const uint16_t *g, const uint16_t *wt, const uint16_t *act,
                               uint16_t *gin, void *stream) {
    if (!img_ok(n_img, H, W) || !g || !wt || !act || !gin) {
        set_error("cmm_cnn_backward_data_relu: bad argument (width a multiple of 32)");
        return CMM_ERR_ARG;
    }
    return conv3(n_img, H, W, g, wt, nullptr, 3, gin, act, stream);
}

int cmm_cnn_backward_weights(int n_img, int H, int W, const uint16_t *g, const uint16_t *in, float *dw, float *db,
                             void *stream) {
    if (!img_ok(n_img, H, W) || !g || !in || !dw || !db) {
        set_error("cmm_cnn_backward_weights: bad argument");
        return CMM_ERR_ARG;
    }
    // at most 2 blocks per CU: each block adds its sums with 9248 FP32 atomics
    const int bands = (H + kWgRows - 1) / kWgRows;
    const long units = (long)n_img * bands * (W / 32);
    const long blocks = std::min<long>((units + kWgWaves - 1) / kWgWaves, 512);
    hipLaunchKernelGGL(k_conv3_wgrad, dim3((unsigned)blocks), dim3(64 * kWgWaves), 0, S(stream), g, in, dw, db, H, W,
                       bands, units);
    return lc("k_conv3_wgrad");
}

int cmm_cnn_relu_mask(long n, const uint16_t *act, uint16_t *g, void *stream) {
    if (n < 0 || (n > 0 && (!act || !g))) { set_error("cmm_cnn_relu_mask: bad argument"); return CMM_ERR_ARG; }
    if (n == 0) return CMM_OK;
    hipLaunchKernelGGL(k_relu_mask, dim3(blocks(n)), dim3(256), 0, S(stream), n, act, g);
    return lc("k_relu_mask");
}

int cmm_cnn_fill(long n, float *p, float v, void *stream) {
    if (n < 0 || (n > 0 && !p)) { set_error("cmm_cnn_fill: bad argument"); return CMM_ERR_ARG; }
    if (n == 0) return CMM_OK;
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks(n)), dim3(256), 0, S(stream), n, p, v);
    return lc("k_fill_f32");
}

int cmm_cnn_weights_bf16(int n_layers, const float *w32, uint16_t *w, uint16_t *wt, void *stream) {
    if (n_layers <= 0 || !w32 || !w || !wt) { set_error("cmm_cnn_weights_bf16: bad argument"); return CMM_ERR_ARG; }
    const long n = (long)n_layers * kCh * 9 * kCh;
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks(n)), dim3(256), 0, S(stream), n, w32, w, wt);
    return lc("k_weights_bf16");
}

int cmm_cnn_adamw(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                  float wd, float gscale, int step, void *stream) {
    if (n <= 0 || !p || !g || !m || !v || step < 1) { set_error("cmm_cnn_adamw: bad argument"); return CMM_ERR_ARG; }
    const float bc1 = 1.f - std::pow(b1, (float)step), bc2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adamw, dim3(blocks(n)), dim3(256), 0, S(stream), n, p, g, m, v, lr, b1, b2, eps, wd, gscale,
                       bc1, bc2);
    return lc("k_adamw");
}

int cmm_cnn_input(long n, const double *psi0, const double *psi1, const double *psi2, const float *scales,
                  uint16_t *out, void *stream) {
    if (n <= 0 || !psi0 || !psi1 || !psi2 || !scales || !out) { set_error("cmm_cnn_input: bad argument"); return CMM_ERR_ARG; }
    hipLaunchKernelGGL(k_cnn_input, dim3(blocks(n)), dim3(256), 0, S(stream), n, psi0, psi1, psi2, scales, out);
    return lc("k_cnn_input");
}

int cmm_cnn_target(long n, const double *psi0, const double *psi1, const double *psi2, const double *psi_next,
                   const double *psi_half, const double *L1, const double *L12, const float *rscales, float *target,
                   void *stream) {
    if (n <= 0 || !psi0 || !psi1 || !psi2 || !psi_next || !psi_half || !L1 || !L12 || !rscales || !target) {
        set_error("cmm_cnn_target: bad argument");
        return CMM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_cnn_target, dim3(blocks(n)), dim3(256), 0, S(stream), n, psi0, psi1, psi2, psi_next, psi_half,
                       L1[0], L1[1], L1[2], L12[0], L12[1], L12[2], rscales, target);
    return lc("k_cnn_target");
}

int cmm_cnn_mse(long n, const float *pred, const float *target, uint16_t *g, double *loss, void *stream) {
    if (n <= 0 || !pred || !target || !g || !loss) { set_error("cmm_cnn_mse: bad argument"); return CMM_ERR_ARG; }
    hipLaunchKernelGGL(k_cnn_mse, dim3((unsigned)std::min<long>(blocks(n), 4096)), dim3(256), 0, S(stream), n, pred,
                       target, g, loss);
    return lc("k_cnn_mse");
}

int cmm_cnn_correction(long n, const float *pred, const float *rscales, double *d1, double *d12, void *stream) {
    if (n <= 0 || !pred || !rscales || !d1 || !d12) { set_error("cmm_cnn_correction: bad argument"); return CMM_ERR_ARG; }
    hipLaunchKernelGGL(k_cnn_correction, dim3(blocks(n)), dim3(256), 0, S(stream), n, pred, rscales, d1, d12);
    return lc("k_cnn_correction");
}

int cmm_extrapolate(long n, const double *psi0, const double *psi1, const double *psi2, const double *w,
                    const double *corr, double *out, void *stream) {
    if (n <= 0 || !psi0 || !psi1 || !psi2 || !w || !out) { set_error("cmm_extrapolate: bad argument"); return CMM_ERR_ARG; }
    hipLaunchKernelGGL(k_extrapolate, dim3(blocks(n)), dim3(256), 0, S(stream), n, psi0, psi1, psi2, w[0], w[1], w[2],
                       corr, out);
    return lc("k_extrapolate");
}

int cmm_dev_norms(long n, const double *a, const double *b, double *out2, void *stream) {
    if (n <= 0 || !a || !b || !out2) { set_error("cmm_dev_norms: bad argument"); return CMM_ERR_ARG; }
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, S(stream), out2);
    hipLaunchKernelGGL(k_dev_norms, dim3((unsigned)std::min<long>(1024, blocks(n))), dim3(256), 0, S(stream), n, a, b,
                       out2);
    return lc("k_dev_norms");
}

}  // extern "C"
