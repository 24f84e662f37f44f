// Kernels of the learned warm start: a 3-level 3D convolutional U-Net
// trained on the GPU.  No reference counterpart.
//
// Layout: activations are channels-last, [N][D][H][W][C] with C a multiple
// of 8, stored as BF16 (raw 16-bit words); gradients and master weights are
// FP32.  A 3x3x3 convolution (zero padding 1) is an implicit GEMM
//     out[p][co] = sum_{k = tap*Ci + ci} W[co][k] * in[p + off(tap)][ci]
// computed with gfx950's v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation): the weights are the A operand (rows = output channels),
// the gathered activations the B operand (columns = 16 positions), so each
// lane ends with 4 consecutive output channels of one position and stores
// them as one 8-byte word.  Weights are [Co][Kpad] with Kpad = 27*Ci rounded
// up to 32 (zero-filled).
//
// The backward data pass is the same kernel applied to the gradient with
// the weights transposed and the taps mirrored (conv_wt_transpose); the
// weight gradient is a second MFMA GEMM reducing over positions with float
// atomics into FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/stencil_nn.h"
#include "st_device.hpp"
#include "st_internal.hpp"

namespace st {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f)
{
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((x >> 16) | 0x40);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (bf16_t)(x >> 16);
}

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const bf16_t *p)
{
    const u32x4 v = *reinterpret_cast<const u32x4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 zero8()
{
    const u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, z);
}

struct ConvGeom {
    int N, D, H, W;
    int Ci, Co;
    int Kpad;       // row length of the weight matrix
    int P;          // N*D*H*W (below 2^31)
};

__device__ __forceinline__ void decode(const ConvGeom &g, int p, int &n, int &z, int &y, int &x)
{
    x = p % g.W;
    int t = p / g.W;
    y = t % g.H;
    t /= g.H;
    z = t % g.D;
    n = t / g.D;
}

__device__ __forceinline__ bf16x8 pack8(const unsigned short (&v)[8])
{
    u32x4 w;
    w[0] = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    w[1] = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    w[2] = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
    w[3] = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
    return __builtin_bit_cast(bf16x8, w);
}

// ---------------------------------------------------------------- conv forward / backward data

// Wave tile: MB blocks of 16 positions x NB blocks of 16 output channels.
template <int NB, int MB>
__global__ __launch_bounds__(256) void k_conv3(const bf16_t *__restrict__ act,
                                               const bf16_t *__restrict__ wt,
                                               const float *__restrict__ bias,
                                               bf16_t *__restrict__ out_bf,
                                               float *__restrict__ out_f, ConvGeom g, int relu,
                                               int accumulate)
{
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, grp = lane >> 4;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nct = (g.Co + 16 * NB - 1) / (16 * NB);
    const long long pt = wave / nct;
    const int ct = (int)(wave % nct);
    const int p0 = (int)(pt * 16 * MB);
    const int c0 = ct * 16 * NB;
    if (pt * 16 * MB >= g.P) return;

    int pn[MB], pz[MB], py[MB], px[MB];
    bool pv[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int p = p0 + mb * 16 + j;
        pv[mb] = p < g.P;
        decode(g, pv[mb] ? p : 0, pn[mb], pz[mb], py[mb], px[mb]);
    }
    f32x4 acc[NB][MB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[nb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nk = g.Kpad / 32;
    for (int kc = 0; kc < nk; ++kc) {
        const int k = kc * 32 + grp * 8;
        const int tap = k / g.Ci;
        const int ci = k - tap * g.Ci;
        bf16x8 a[NB], b[MB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int co = c0 + nb * 16 + j;
            a[nb] = co < g.Co ? load8(wt + (long long)co * g.Kpad + k) : zero8();
        }
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int zz = pz[mb] + dz, yy = py[mb] + dy, xx = px[mb] + dx;
            const bool ok = pv[mb] && tap < 27 && zz >= 0 && zz < g.D && yy >= 0 && yy < g.H &&
                            xx >= 0 && xx < g.W;
            const long long q = (((long long)pn[mb] * g.D + zz) * g.H + yy) * g.W + xx;
            b[mb] = ok ? load8(act + q * g.Ci + ci) : zero8();
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                acc[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[nb], b[mb], acc[nb][mb], 0, 0, 0);
    }
    // lane holds D[co = 4*grp + v][position j] of each (nb, mb) block
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int cb = c0 + nb * 16 + 4 * grp;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            if (!pv[mb]) continue;
            const long long p = (long long)p0 + mb * 16 + j;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float t = acc[nb][mb][e];
                if (bias && cb + e < g.Co) t += bias[cb + e];
                if (relu) t = t > 0.f ? t : 0.f;
                v[e] = t;
            }
            if (out_f) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (cb + e < g.Co) {
                        float *o = out_f + p * g.Co + cb + e;
                        *o = accumulate ? *o + v[e] : v[e];
                    }
            }
            if (out_bf) {
                if (cb + 3 < g.Co) {
                    const uint32_t lo = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                    const uint32_t hi = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                    *reinterpret_cast<uint2 *>(out_bf + p * g.Co + cb) = make_uint2(lo, hi);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (cb + e < g.Co) out_bf[p * g.Co + cb + e] = f2bf(v[e]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- weight gradient

// dW[co][k] += sum_p dY[p][co] * X[p + off(tap(k))][ci(k)]
// MFMA rows = output channels (A = dY^T), columns = 16 weight columns k
// (B = shifted X), reduction over 32 positions per instruction; each wave
// covers `pchunk` positions and adds its partial sums atomically.
template <int NB>
__global__ __launch_bounds__(256) void k_conv3_wgrad(const bf16_t *__restrict__ X,
                                                     const bf16_t *__restrict__ dY,
                                                     float *__restrict__ dW, ConvGeom g,
                                                     int pchunk)
{
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, grp = lane >> 4;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int K = 27 * g.Ci;
    const int nkt = (K + 16 * NB - 1) / (16 * NB);
    const int nct = (g.Co + 15) / 16;
    const long long nchunks = ((long long)g.P + pchunk - 1) / pchunk;
    const long long total = nchunks * nkt * nct;
    if (wave >= total) return;
    const int kt = (int)(wave % nkt);
    const long long r2 = wave / nkt;
    const int ct = (int)(r2 % nct);
    const int pc = (int)(r2 / nct);
    const int co = ct * 16 + j;  // A-operand row of this lane
    int ci[NB], kk[NB], dz[NB], dy[NB], dx[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        kk[nb] = kt * 16 * NB + nb * 16 + j;  // B-operand column of this lane
        const int kc = kk[nb] < K ? kk[nb] : 0;
        const int tap = kc / g.Ci;
        ci[nb] = kc - tap * g.Ci;
        dz[nb] = tap / 9 - 1;
        dy[nb] = (tap / 3) % 3 - 1;
        dx[nb] = tap % 3 - 1;
    }
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int pbeg = pc * pchunk, pend = min(pbeg + pchunk, g.P);
    for (int ps = pbeg; ps < pend; ps += 32) {
        // the 8 positions of this lane group, walked from one decode
        int n0, z0, y0, x0;
        const int pfirst = ps + grp * 8;
        decode(g, min(pfirst, g.P - 1), n0, z0, y0, x0);
        int pn[8], pz[8], py[8], px[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            pn[e] = n0; pz[e] = z0; py[e] = y0; px[e] = x0;
            if (++x0 == g.W) {
                x0 = 0;
                if (++y0 == g.H) {
                    y0 = 0;
                    if (++z0 == g.D) { z0 = 0; ++n0; }
                }
            }
        }
        unsigned short av[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int p = pfirst + e;
            av[e] = (p < pend && co < g.Co) ? dY[(long long)p * g.Co + co] : (unsigned short)0;
        }
        const bf16x8 a = pack8(av);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            unsigned short bv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int zz = pz[e] + dz[nb], yy = py[e] + dy[nb], xx = px[e] + dx[nb];
                const bool ok = pfirst + e < pend && kk[nb] < K && zz >= 0 && zz < g.D && yy >= 0 &&
                                yy < g.H && xx >= 0 && xx < g.W;
                bv[e] = ok ? X[((((long long)pn[e] * g.D + zz) * g.H + yy) * g.W + xx) * g.Ci + ci[nb]]
                           : (unsigned short)0;
            }
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pack8(bv), acc[nb], 0, 0, 0);
        }
    }
    // lane holds D[row co = ct*16 + 4*grp + v][column k = kt*16*NB + nb*16 + j]
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        if (kk[nb] >= K) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = ct * 16 + 4 * grp + e;
            if (c < g.Co) atomicAdd(dW + (long long)c * g.Kpad + kk[nb], acc[nb][e]);
        }
    }
}

// ---------------------------------------------------------------- weight gradient, channel-major operands

// The MFMA operands of the weight gradient run along positions (8
// consecutive positions of one channel per lane), which channels-last
// activations hold Ci values apart: the gather above issues one 2-byte load
// per value.  When rows are a multiple of 8 positions, both operands are
// first transposed to channel-major [C][P] copies, and then every operand is
// one 16-byte load: the dY rows directly, the X rows of a tap with dx = -1
// or +1 from two aligned loads joined by a 16-bit funnel shift (a group of 8
// positions never crosses a row, so only its first or last value can fall
// outside the row, and it is zeroed).

// [P][C] -> [C][P] (16-bit words) through LDS, 64 positions per workgroup.
constexpr int kTrMaxC = 128;
__global__ __launch_bounds__(256) void k_pc_transpose(const bf16_t *__restrict__ src, bf16_t *__restrict__ dst,
                                                      int P, int C)
{
    __shared__ bf16_t t[kTrMaxC * 66];
    const long long p0 = (long long)blockIdx.x * 64;
    const int np = (int)min(64LL, (long long)P - p0);
    for (int i = threadIdx.x; i < np * C; i += 256) {
        const int pp = i / C, c = i - pp * C;
        t[c * 66 + pp] = src[p0 * C + i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * C; i += 256) {
        const int c = i >> 6, pp = i & 63;
        if (pp < np) dst[(long long)c * P + p0 + pp] = t[c * 66 + pp];
    }
}

__device__ __forceinline__ u32x4 load8u(const bf16_t *p) { return *reinterpret_cast<const u32x4 *>(p); }

// dW[co][k] += sum_p dYt[co][p] * Xt[ci(k)][p + off(tap(k))]; the same
// tiling as k_conv3_wgrad (rows = output channels, 16*NB weight columns,
// 32 positions per MFMA, `pchunk` positions per wave, atomics at the end).
// Requires W % 8 == 0 and pchunk % 32 == 0.
template <int NB>
__global__ __launch_bounds__(256) void k_conv3_wgrad_t(const bf16_t *__restrict__ Xt,
                                                       const bf16_t *__restrict__ dYt,
                                                       float *__restrict__ dW, ConvGeom g, int pchunk)
{
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, grp = lane >> 4;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int K = 27 * g.Ci;
    const int nkt = (K + 16 * NB - 1) / (16 * NB);
    const int nct = (g.Co + 15) / 16;
    const long long nchunks = ((long long)g.P + pchunk - 1) / pchunk;
    const long long total = nchunks * nkt * nct;
    if (wave >= total) return;
    const int kt = (int)(wave % nkt);
    const long long r2 = wave / nkt;
    const int ct = (int)(r2 % nct);
    const int pc = (int)(r2 / nct);
    const int co = ct * 16 + j;  // A-operand row of this lane
    const long long HW = (long long)g.H * g.W;
    const bf16_t *arow = dYt + (long long)min(co, g.Co - 1) * g.P;
    int kk[NB], dz[NB], dy[NB], dx[NB];
    const bf16_t *brow[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        kk[nb] = kt * 16 * NB + nb * 16 + j;  // B-operand column of this lane
        const int kc = kk[nb] < K ? kk[nb] : 0;
        const int tap = kc / g.Ci;
        const int ci = kc - tap * g.Ci;
        dz[nb] = tap / 9 - 1;
        dy[nb] = (tap / 3) % 3 - 1;
        dx[nb] = tap % 3 - 1;
        brow[nb] = Xt + (long long)ci * g.P;
    }
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int pbeg = pc * pchunk, pend = min(pbeg + pchunk, g.P);
    const u32x4 zero = {0u, 0u, 0u, 0u};
    for (int ps = pbeg; ps < pend; ps += 32) {
        // this lane group's 8 positions: one row segment (W % 8 == 0)
        const int pf = ps + grp * 8;
        const bool pin = pf < pend;
        int n0, z0, y0, x0;
        decode(g, min(pf, g.P - 8), n0, z0, y0, x0);
        const u32x4 av = (pin && co < g.Co) ? load8u(arow + pf) : zero;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int zz = z0 + dz[nb], yy = y0 + dy[nb];
            const bool ok = pin && kk[nb] < K && zz >= 0 && zz < g.D && yy >= 0 && yy < g.H;
            u32x4 b = zero;
            if (ok) {
                const bf16_t *q = brow[nb] + pf + (long long)dz[nb] * HW + (long long)dy[nb] * g.W;
                const u32x4 v = load8u(q);
                if (dx[nb] == 0) {
                    b = v;
                } else if (dx[nb] < 0) {
                    // values x0-1 .. x0+6: the last of the group before, then v[0..6]
                    const unsigned prev = x0 > 0 ? load8u(q - 8)[3] : 0u;
                    b[0] = __builtin_amdgcn_alignbit(v[0], prev, 16);
                    b[1] = __builtin_amdgcn_alignbit(v[1], v[0], 16);
                    b[2] = __builtin_amdgcn_alignbit(v[2], v[1], 16);
                    b[3] = __builtin_amdgcn_alignbit(v[3], v[2], 16);
                } else {
                    // values x0+1 .. x0+8: v[1..7], then the first of the group after
                    const unsigned next = x0 + 8 < g.W ? load8u(q + 8)[0] : 0u;
                    b[0] = __builtin_amdgcn_alignbit(v[1], v[0], 16);
                    b[1] = __builtin_amdgcn_alignbit(v[2], v[1], 16);
                    b[2] = __builtin_amdgcn_alignbit(v[3], v[2], 16);
                    b[3] = __builtin_amdgcn_alignbit(next, v[3], 16);
                }
            }
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                              __builtin_bit_cast(bf16x8, b), acc[nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        if (kk[nb] >= K) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = ct * 16 + 4 * grp + e;
            if (c < g.Co) atomicAdd(dW + (long long)c * g.Kpad + kk[nb], acc[nb][e]);
        }
    }
}

// ---------------------------------------------------------------- small layers

// W[Co][Kpad] (k = tap*Ci + ci)  ->  Wt[Ci][KpadT] (k' = tap'*Co + co) with
// tap' = 26 - tap: the backward-data weights.
__global__ void k_wt_transpose(const bf16_t *__restrict__ w, bf16_t *__restrict__ wt, int Ci, int Co,
                               int Kpad, int KpadT)
{
    const long long n = (long long)Ci * KpadT;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(t / KpadT);
        const int k2 = (int)(t % KpadT);
        bf16_t v = 0;
        if (k2 < 27 * Co) {
            const int tap2 = k2 / Co, co = k2 % Co;
            v = w[(long long)co * Kpad + (26 - tap2) * Ci + ci];
        }
        wt[t] = v;
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ a, bf16_t *__restrict__ b, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        b[t] = f2bf(a[t]);
}

__global__ void k_bf16_to_f32(const bf16_t *__restrict__ a, float *__restrict__ b, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        b[t] = bf2f(a[t]);
}

// g_out(bf16) = g(f32) * (y > 0)   (ReLU backward, y = the layer's BF16 output)
__global__ void k_relu_bwd(const float *__restrict__ g, const bf16_t *__restrict__ y,
                           bf16_t *__restrict__ out, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        out[t] = bf2f(y[t]) > 0.f ? f2bf(g[t]) : (bf16_t)0;
}

// db[co] += sum_p g[p][co]: each thread sums one channel over a strided
// set of positions, the block reduces its 256/C partial sums per channel in
// LDS, and one atomic per channel and block goes out (an atomic per thread
// put every thread of the grid on the same C addresses: 1.6 ms per call).
__global__ __launch_bounds__(256) void k_bias_grad(const bf16_t *__restrict__ g, float *__restrict__ db,
                                                   long long P, int C)
{
    __shared__ float red[256];
    const int c = threadIdx.x % C;  // C is a power of two dividing 256
    const int lanes = 256 / C;
    float s = 0.f;
    for (long long p = blockIdx.x * (long long)lanes + threadIdx.x / C; p < P;
         p += (long long)gridDim.x * lanes)
        s += bf2f(g[p * C + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    // entries t and t + stride hold the same channel while stride >= C
    for (int stride = 128; stride >= C; stride >>= 1) {
        if ((int)threadIdx.x < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
    if ((int)threadIdx.x < C) atomicAdd(db + threadIdx.x, red[threadIdx.x]);
}

// 2x2x2 average pool, channels-last.  in [N][2D][2H][2W][C] -> out [N][D][H][W][C]
__global__ void k_pool_fwd(const bf16_t *__restrict__ in, bf16_t *__restrict__ out, int N, int D,
                           int H, int W, int C)
{
    const long long n = (long long)N * D * H * W * C;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C);
        long long p = t / C;
        const int x = (int)(p % W); p /= W;
        const int y = (int)(p % H); p /= H;
        const int z = (int)(p % D);
        const int b = (int)(p / D);
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
            s += bf2f(in[((((long long)b * 2 * D + zz) * 2 * H + yy) * 2 * W + xx) * C + c]);
        }
        out[t] = f2bf(s * 0.125f);
    }
}

// backward of the pool: gin (f32, [N][2D][2H][2W][C]) = gout / 8 (accumulate)
__global__ void k_pool_bwd(const float *__restrict__ gout, float *__restrict__ gin, int N, int D,
                           int H, int W, int C)
{
    const long long n = (long long)N * 2 * D * 2 * H * 2 * W * C;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C);
        long long p = t / C;
        const int x = (int)(p % (2 * W)); p /= 2 * W;
        const int y = (int)(p % (2 * H)); p /= 2 * H;
        const int z = (int)(p % (2 * D));
        const int b = (int)(p / (2 * D));
        gin[t] += 0.125f * gout[((((long long)b * D + z / 2) * H + y / 2) * W + x / 2) * C + c];
    }
}

// nearest x2 upsample into channels [c0, c0+C) of a Ct-channel tensor:
// out[N][2D][2H][2W][Ct]; in [N][D][H][W][C]
__global__ void k_up_fwd(const bf16_t *__restrict__ in, bf16_t *__restrict__ out, int N, int D, int H,
                         int W, int C, int Ct, int c0)
{
    const long long n = (long long)N * 2 * D * 2 * H * 2 * W * C;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C);
        long long p = t / C;
        const int x = (int)(p % (2 * W)); p /= 2 * W;
        const int y = (int)(p % (2 * H)); p /= 2 * H;
        const int z = (int)(p % (2 * D));
        const int b = (int)(p / (2 * D));
        const long long po = (((long long)b * 2 * D + z) * 2 * H + y) * 2 * W + x;
        out[po * Ct + c0 + c] = in[((((long long)b * D + z / 2) * H + y / 2) * W + x / 2) * C + c];
    }
}

// backward of the upsample: gin[N][D][H][W][C] = sum of the 8 children of
// channel c0+c in gout[N][2D][2H][2W][Ct]
__global__ void k_up_bwd(const float *__restrict__ gout, float *__restrict__ gin, int N, int D, int H,
                         int W, int C, int Ct, int c0)
{
    const long long n = (long long)N * D * H * W * C;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t % C);
        long long p = t / C;
        const int x = (int)(p % W); p /= W;
        const int y = (int)(p % H); p /= H;
        const int z = (int)(p % D);
        const int b = (int)(p / D);
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
            s += gout[((((long long)b * 2 * D + zz) * 2 * H + yy) * 2 * W + xx) * Ct + c0 + c];
        }
        gin[t] = s;
    }
}

// copy channels: dst[p][dc0 + c] = src[p][sc0 + c], c < C  (concat forward,
// gradient split backward); T is bf16_t or float
template <typename T>
__global__ void k_chan_copy(const T *__restrict__ src, T *__restrict__ dst, long long P, int C,
                            int Cs, int sc0, int Cd, int dc0, int accumulate)
{
    const long long n = P * C;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const long long p = t / C;
        const int c = (int)(t % C);
        const T v = src[p * Cs + sc0 + c];
        if constexpr (sizeof(T) == 4) {
            if (accumulate) dst[p * Cd + dc0 + c] += v;
            else dst[p * Cd + dc0 + c] = v;
        } else {
            dst[p * Cd + dc0 + c] = v;
        }
    }
}

// 1x1 head: y[p] = b + sum_c w[c] * x[p][c]   (x BF16, w/b/y FP32)
__global__ void k_head_fwd(const bf16_t *__restrict__ x, const float *__restrict__ w,
                           const float *__restrict__ b, float *__restrict__ y, long long P, int C)
{
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        float s = b[0];
        for (int c = 0; c < C; ++c) s += w[c] * bf2f(x[p * C + c]);
        y[p] = s;
    }
}

// head backward: gx[p][c] = g[p] * w[c] (f32); dw[c] += sum_p g[p] x[p][c]; db += sum_p g[p]
template <int C>
__global__ __launch_bounds__(256) void k_head_bwd(const bf16_t *__restrict__ x,
                                                  const float *__restrict__ w,
                                                  const float *__restrict__ g,
                                                  float *__restrict__ gx, float *__restrict__ dw,
                                                  float *__restrict__ db, long long P)
{
    __shared__ float red[C + 1];
    if ((int)threadIdx.x <= C) red[threadIdx.x] = 0.f;
    __syncthreads();
    float sdw[C];
#pragma unroll
    for (int c = 0; c < C; ++c) sdw[c] = 0.f;
    float sdb = 0.f;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const float gp = g[p];
        sdb += gp;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            gx[p * C + c] = gp * w[c];
            sdw[c] += gp * bf2f(x[p * C + c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) atomicAdd(&red[c], sdw[c]);
    atomicAdd(&red[C], sdb);
    __syncthreads();
    if ((int)threadIdx.x < C) atomicAdd(dw + threadIdx.x, red[threadIdx.x]);
    if ((int)threadIdx.x == C) atomicAdd(db, red[C]);
}

// prediction u = base + y on the interior of each n^3 sample, base elsewhere;
// loss += sum (u - target)^2 over the interior; g = 2 (u - target) / count
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ base,
                                             const float *__restrict__ y,
                                             const float *__restrict__ target,
                                             float *__restrict__ g, double *__restrict__ loss,
                                             float *__restrict__ pred, int N, int n, double scale)
{
    __shared__ double red[256];
    const long long P = (long long)N * n * n * n;
    double acc = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(p % n), yy = (int)((p / n) % n), z = (int)((p / ((long long)n * n)) % n);
        const bool in = x > 0 && x < n - 1 && yy > 0 && yy < n - 1 && z > 0 && z < n - 1;
        const float u = in ? base[p] + y[p] : base[p];
        if (pred) pred[p] = u;
        float gg = 0.f;
        if (in && target) {
            const float d = u - target[p];
            acc += (double)d * d;
            gg = (float)(2.0 * scale) * d;
        }
        if (g) g[p] = gg;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

// out = a*x + b*y (FP32)
__global__ void k_axpby(float *__restrict__ out, float a, const float *__restrict__ x, float b,
                        const float *__restrict__ y, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        out[t] = a * x[t] + b * y[t];
}

// loss += sum x^2 (FP64), g = 2*scale*x
__global__ __launch_bounds__(256) void k_sumsq(const float *__restrict__ x, float *__restrict__ g,
                                               double *__restrict__ loss, long long n, double scale)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const float v = x[t];
        acc += (double)v * v;
        if (g) g[t] = (float)(2.0 * scale) * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

// Adam (Kingma & Ba 2015) with bias correction; writes the BF16 copy of the
// weights for the MFMA kernels (zero columns of padded rows stay zero
// because their gradients are zero).
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, bf16_t *__restrict__ wbf, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const float gt = g[t];
        const float mt = b1 * m[t] + (1.f - b1) * gt;
        const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
        m[t] = mt;
        v[t] = vt;
        const float upd = lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        const float wt = w[t] - upd;
        w[t] = wt;
        if (wbf) wbf[t] = f2bf(wt);
    }
}

// Input features of an n^3 sample from its field u (boundary shell filled):
// channel 0 = transfinite blend of the six faces
//   (1/3) sum over axes [(1-s) F_lo + s F_hi],  s = coordinate / (n-1),
// channel 1 = the boundary value on the shell (0 inside), channels 2..C-1 = 0.
// base[p] receives the blend inside and the boundary value on the shell.
__global__ void k_features(const float *__restrict__ u, bf16_t *__restrict__ feat,
                           float *__restrict__ base, int N, int n, int C)
{
    const long long P = (long long)N * n * n * n;
    const long long nn = (long long)n * n;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(p % n), y = (int)((p / n) % n), z = (int)((p / nn) % n);
        const long long s0 = p - (long long)z * nn - (long long)y * n - x;  // sample origin
        const float inv = 1.0f / (float)(n - 1);
        const float sx = x * inv, sy = y * inv, sz = z * inv;
        const float fx = (1.f - sx) * u[s0 + (long long)z * nn + (long long)y * n] +
                         sx * u[s0 + (long long)z * nn + (long long)y * n + (n - 1)];
        const float fy = (1.f - sy) * u[s0 + (long long)z * nn + x] +
                         sy * u[s0 + (long long)z * nn + (long long)(n - 1) * n + x];
        const float fz = (1.f - sz) * u[s0 + (long long)y * n + x] +
                         sz * u[s0 + (long long)(n - 1) * nn + (long long)y * n + x];
        const float blend = (fx + fy + fz) * (1.0f / 3.0f);
        const bool shell = x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
        const float bv = shell ? u[p] : 0.f;
        // channels: blend, shell value, the three coordinates, the distance
        // to the nearest face (scaled to [0, 1]), zeros
        const float dist =
            2.f * fminf(fminf(fminf(sx, 1.f - sx), fminf(sy, 1.f - sy)), fminf(sz, 1.f - sz));
        for (int c = 0; c < C; ++c) {
            const float v = c == 0 ? blend
                          : c == 1 ? bv
                          : c == 2 ? sx
                          : c == 3 ? sy
                          : c == 4 ? sz
                          : c == 5 ? dist
                          : 0.f;
            feat[p * C + c] = f2bf(v);
        }
        if (base) base[p] = shell ? u[p] : blend;
    }
}

// Trilinear interpolation of the interior of a node-centred m^3 field onto
// an M^3 field: fine node I maps to coarse coordinate I*(m-1)/(M-1).  The
// shell of `out` is left untouched (it holds the exact boundary).
__global__ void k_trilinear(const float *__restrict__ in, float *__restrict__ out, int m, int M)
{
    const long long n = (long long)M * M * M;
    const float sc = (float)(m - 1) / (float)(M - 1);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int X = (int)(t % M), Y = (int)((t / M) % M), Z = (int)(t / ((long long)M * M));
        if (X == 0 || Y == 0 || Z == 0 || X == M - 1 || Y == M - 1 || Z == M - 1) continue;
        const float cx = X * sc, cy = Y * sc, cz = Z * sc;
        const int x0 = min((int)cx, m - 2), y0 = min((int)cy, m - 2), z0 = min((int)cz, m - 2);
        const float fx = cx - x0, fy = cy - y0, fz = cz - z0;
        auto at = [&](int a, int b, int c) { return in[((long long)c * m + b) * m + a]; };
        const float c00 = at(x0, y0, z0) * (1 - fx) + at(x0 + 1, y0, z0) * fx;
        const float c10 = at(x0, y0 + 1, z0) * (1 - fx) + at(x0 + 1, y0 + 1, z0) * fx;
        const float c01 = at(x0, y0, z0 + 1) * (1 - fx) + at(x0 + 1, y0, z0 + 1) * fx;
        const float c11 = at(x0, y0 + 1, z0 + 1) * (1 - fx) + at(x0 + 1, y0 + 1, z0 + 1) * fx;
        const float c0 = c00 * (1 - fy) + c10 * fy, c1 = c01 * (1 - fy) + c11 * fy;
        out[t] = c0 * (1 - fz) + c1 * fz;
    }
}

// Dirichlet faces from Fourier modes: face value sum_m a_m sin(pi kx s) sin(pi ky t)
// with (s, t) the face's two coordinates scaled to [0, 1]; modes[f*M*3 + m*3 + {0,1,2}]
// = (a, kx, ky) for the six faces x0, x1, y0, y1, z0, z1.  Writes the shell of
// the N samples' n^3 fields (the interior is untouched).
__global__ void k_fourier_faces(float *__restrict__ u, const float *__restrict__ modes, int N, int n,
                                int M)
{
    const long long nn = (long long)n * n;
    const long long P = (long long)N * nn * n;
    const float inv = 1.0f / (float)(n - 1);
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(p % n), y = (int)((p / n) % n), z = (int)((p / nn) % n);
        const int b = (int)(p / (nn * n));
        int f = -1;
        float s = 0.f, t = 0.f;
        if (x == 0 || x == n - 1) { f = x == 0 ? 0 : 1; s = y * inv; t = z * inv; }
        else if (y == 0 || y == n - 1) { f = y == 0 ? 2 : 3; s = x * inv; t = z * inv; }
        else if (z == 0 || z == n - 1) { f = z == 0 ? 4 : 5; s = x * inv; t = y * inv; }
        if (f < 0) continue;
        const float *md = modes + ((long long)b * 6 + f) * M * 3;
        float v = 0.f;
        for (int m = 0; m < M; ++m)
            v += md[3 * m] * sinpif(md[3 * m + 1] * s) * sinpif(md[3 * m + 2] * t);
        u[p] = v;
    }
}

}  // namespace st

// ---------------------------------------------------------------- C ABI

using namespace st;

static dim3 gs_grid(long long n)
{
    long long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

static ConvGeom make_conv(int N, int D, int H, int W, int Ci, int Co, int Kpad)
{
    ConvGeom g;
    g.N = N; g.D = D; g.H = H; g.W = W; g.Ci = Ci; g.Co = Co; g.Kpad = Kpad;
    g.P = (int)((long long)N * D * H * W);
    return g;
}

extern "C" {

int st_nn_kpad(int ci) { return (27 * ci + 31) / 32 * 32; }

st_status st_nn_conv3(const void *act, const void *w, const float *bias, void *out_bf16, float *out_f32,
                      int N, int D, int H, int W, int Ci, int Co, int relu, int accumulate, void *stream)
{
    if (Ci % 8 || Co % 4 || Co < 4 || N < 1 || D < 1 || H < 1 || W < 1)
        return fail(ST_EINVAL, "conv3: Ci must be a multiple of 8 and Co of 4 (got %d, %d)", Ci, Co);
    if ((long long)N * D * H * W >= (1LL << 31)) return fail(ST_EINVAL, "conv3: too many positions");
    const ConvGeom g = make_conv(N, D, H, W, Ci, Co, st_nn_kpad(Ci));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = Co >= 64 ? 4 : (Co >= 32 ? 2 : 1);
    const long long waves = ((g.P + 63) / 64) * ((Co + 16 * nb - 1) / (16 * nb));
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    const bf16_t *a = static_cast<const bf16_t *>(act);
    const bf16_t *wt = static_cast<const bf16_t *>(w);
    bf16_t *ob = static_cast<bf16_t *>(out_bf16);
    if (nb == 4) hipLaunchKernelGGL((k_conv3<4, 4>), grid, block, 0, s, a, wt, bias, ob, out_f32, g, relu, accumulate);
    else if (nb == 2) hipLaunchKernelGGL((k_conv3<2, 4>), grid, block, 0, s, a, wt, bias, ob, out_f32, g, relu, accumulate);
    else hipLaunchKernelGGL((k_conv3<1, 4>), grid, block, 0, s, a, wt, bias, ob, out_f32, g, relu, accumulate);
    return hip_check(hipGetLastError(), "conv3 launch");
}

st_status st_nn_conv3_wgrad(const void *x, const void *dy, float *dw, int N, int D, int H, int W, int Ci,
                            int Co, void *stream)
{
    if (Ci % 8) return fail(ST_EINVAL, "conv3 wgrad: Ci must be a multiple of 8");
    if ((long long)N * D * H * W >= (1LL << 31)) return fail(ST_EINVAL, "conv3 wgrad: too many positions");
    const ConvGeom g = make_conv(N, D, H, W, Ci, Co, st_nn_kpad(Ci));
    const int pchunk = 2048;
    const int K = 27 * Ci;
    const long long nkt = (K + 63) / 64, nct = (Co + 15) / 16, nch = (g.P + pchunk - 1) / pchunk;
    const long long waves = nkt * nct * nch;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)((waves + 3) / 4));
    if (W % 8 == 0 && Ci <= kTrMaxC && Co <= kTrMaxC) {
        // channel-major copies of both operands in the workspace (slot 0)
        void *ws = nullptr;
        st_status st = workspace_slot(0, ((size_t)g.P * Ci + (size_t)g.P * Co) * sizeof(bf16_t), s, &ws);
        if (st != ST_OK) return st;
        bf16_t *xt = static_cast<bf16_t *>(ws);
        bf16_t *dyt = xt + (size_t)g.P * Ci;
        const dim3 tg((unsigned)((g.P + 63) / 64));
        hipLaunchKernelGGL(k_pc_transpose, tg, dim3(256), 0, s, static_cast<const bf16_t *>(x), xt, g.P, Ci);
        hipLaunchKernelGGL(k_pc_transpose, tg, dim3(256), 0, s, static_cast<const bf16_t *>(dy), dyt, g.P, Co);
        hipLaunchKernelGGL((k_conv3_wgrad_t<4>), grid, dim3(256), 0, s, xt, dyt, dw, g, pchunk);
        return hip_check(hipGetLastError(), "conv3 wgrad launch");
    }
    hipLaunchKernelGGL((k_conv3_wgrad<4>), grid, dim3(256), 0, s, static_cast<const bf16_t *>(x),
                       static_cast<const bf16_t *>(dy), dw, g, pchunk);
    return hip_check(hipGetLastError(), "conv3 wgrad launch");
}

st_status st_nn_wt_transpose(const void *w, void *wt, int Ci, int Co, void *stream)
{
    const int Kpad = st_nn_kpad(Ci), KpadT = st_nn_kpad(Co);
    hipLaunchKernelGGL(k_wt_transpose, gs_grid((long long)Ci * KpadT), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16_t *>(w),
                       static_cast<bf16_t *>(wt), Ci, Co, Kpad, KpadT);
    return hip_check(hipGetLastError(), "wt transpose launch");
}

st_status st_nn_f32_to_bf16(const float *a, void *b, long long n, void *stream)
{
    hipLaunchKernelGGL(k_f32_to_bf16, gs_grid(n), dim3(256), 0, static_cast<hipStream_t>(stream), a,
                       static_cast<bf16_t *>(b), n);
    return hip_check(hipGetLastError(), "f32->bf16 launch");
}

st_status st_nn_bf16_to_f32(const void *a, float *b, long long n, void *stream)
{
    hipLaunchKernelGGL(k_bf16_to_f32, gs_grid(n), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16_t *>(a), b, n);
    return hip_check(hipGetLastError(), "bf16->f32 launch");
}

st_status st_nn_relu_bwd(const float *g, const void *y, void *out, long long n, void *stream)
{
    hipLaunchKernelGGL(k_relu_bwd, gs_grid(n), dim3(256), 0, static_cast<hipStream_t>(stream), g,
                       static_cast<const bf16_t *>(y), static_cast<bf16_t *>(out), n);
    return hip_check(hipGetLastError(), "relu bwd launch");
}

st_status st_nn_bias_grad(const void *g, float *db, long long P, int C, void *stream)
{
    if (C < 1 || 256 % C || (C & (C - 1))) return fail(ST_EINVAL, "bias grad: C must be a power of two dividing 256");
    long long blocks = (P * C + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16_t *>(g), db, P, C);
    return hip_check(hipGetLastError(), "bias grad launch");
}

st_status st_nn_pool_fwd(const void *in, void *out, int N, int D, int H, int W, int C, void *stream)
{
    hipLaunchKernelGGL(k_pool_fwd, gs_grid((long long)N * D * H * W * C), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16_t *>(in),
                       static_cast<bf16_t *>(out), N, D, H, W, C);
    return hip_check(hipGetLastError(), "pool launch");
}

st_status st_nn_pool_bwd(const float *gout, float *gin, int N, int D, int H, int W, int C, void *stream)
{
    hipLaunchKernelGGL(k_pool_bwd, gs_grid((long long)N * 8 * D * H * W * C), dim3(256), 0,
                       static_cast<hipStream_t>(stream), gout, gin, N, D, H, W, C);
    return hip_check(hipGetLastError(), "pool bwd launch");
}

st_status st_nn_up_fwd(const void *in, void *out, int N, int D, int H, int W, int C, int Ct, int c0,
                       void *stream)
{
    hipLaunchKernelGGL(k_up_fwd, gs_grid((long long)N * 8 * D * H * W * C), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16_t *>(in),
                       static_cast<bf16_t *>(out), N, D, H, W, C, Ct, c0);
    return hip_check(hipGetLastError(), "upsample launch");
}

st_status st_nn_up_bwd(const float *gout, float *gin, int N, int D, int H, int W, int C, int Ct, int c0,
                       void *stream)
{
    hipLaunchKernelGGL(k_up_bwd, gs_grid((long long)N * D * H * W * C), dim3(256), 0,
                       static_cast<hipStream_t>(stream), gout, gin, N, D, H, W, C, Ct, c0);
    return hip_check(hipGetLastError(), "upsample bwd launch");
}

st_status st_nn_chan_copy(const void *src, void *dst, long long P, int C, int Cs, int sc0, int Cd, int dc0,
                          int is_f32, int accumulate, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (is_f32)
        hipLaunchKernelGGL(k_chan_copy<float>, gs_grid(P * C), dim3(256), 0, s,
                           static_cast<const float *>(src), static_cast<float *>(dst), P, C, Cs, sc0,
                           Cd, dc0, accumulate);
    else
        hipLaunchKernelGGL(k_chan_copy<bf16_t>, gs_grid(P * C), dim3(256), 0, s,
                           static_cast<const bf16_t *>(src), static_cast<bf16_t *>(dst), P, C, Cs, sc0,
                           Cd, dc0, 0);
    return hip_check(hipGetLastError(), "channel copy launch");
}

st_status st_nn_head_fwd(const void *x, const float *w, const float *b, float *y, long long P, int C,
                         void *stream)
{
    hipLaunchKernelGGL(k_head_fwd, gs_grid(P), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16_t *>(x), w, b, y, P, C);
    return hip_check(hipGetLastError(), "head launch");
}

st_status st_nn_head_bwd(const void *x, const float *w, const float *g, float *gx, float *dw, float *db,
                         long long P, int C, void *stream)
{
    if (C != 16) return fail(ST_EINVAL, "head: built for 16 channels");
    long long blocks = (P + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_head_bwd<16>, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16_t *>(x), w, g, gx, dw, db, P);
    return hip_check(hipGetLastError(), "head bwd launch");
}

st_status st_nn_mse(const float *base, const float *y, const float *target, float *g, double *loss,
                    float *pred, int N, int n, double scale, void *stream)
{
    long long P = (long long)N * n * n * n;
    long long blocks = (P + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream), base,
                       y, target, g, loss, pred, N, n, scale);
    return hip_check(hipGetLastError(), "mse launch");
}

st_status st_nn_adam(float *w, const float *g, float *m, float *v, void *wbf, long long n, float lr,
                     float b1, float b2, float eps, int step, void *stream)
{
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, gs_grid(n), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v,
                       static_cast<bf16_t *>(wbf), n, lr, b1, b2, eps, c1, c2);
    return hip_check(hipGetLastError(), "adam launch");
}

st_status st_nn_axpby(float *out, float a, const float *x, float b, const float *y, long long n,
                      void *stream)
{
    hipLaunchKernelGGL(k_axpby, gs_grid(n), dim3(256), 0, static_cast<hipStream_t>(stream), out, a, x, b,
                       y, n);
    return hip_check(hipGetLastError(), "axpby launch");
}

st_status st_nn_sumsq(const float *x, float *g, double *loss, long long n, double scale, void *stream)
{
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_sumsq, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                       g, loss, n, scale);
    return hip_check(hipGetLastError(), "sumsq launch");
}

st_status st_nn_features(const float *u, void *feat, float *base, int N, int n, int C, void *stream)
{
    hipLaunchKernelGGL(k_features, gs_grid((long long)N * n * n * n), dim3(256), 0,
                       static_cast<hipStream_t>(stream), u, static_cast<bf16_t *>(feat), base, N, n, C);
    return hip_check(hipGetLastError(), "features launch");
}

st_status st_nn_trilinear(const float *in, float *out, int m, int M, void *stream)
{
    hipLaunchKernelGGL(k_trilinear, gs_grid((long long)M * M * M), dim3(256), 0,
                       static_cast<hipStream_t>(stream), in, out, m, M);
    return hip_check(hipGetLastError(), "trilinear launch");
}

st_status st_nn_fourier_faces(float *u, const float *modes, int N, int n, int M, void *stream)
{
    hipLaunchKernelGGL(k_fourier_faces, gs_grid((long long)N * n * n * n), dim3(256), 0,
                       static_cast<hipStream_t>(stream), u, modes, N, n, M);
    return hip_check(hipGetLastError(), "fourier faces launch");
}

}  // extern "C"
