// SMEM_REGISTER family, row-streaming form: one sweep, radius 1..4, FP32 or
// FP64, whole grid rows per wave.
//
// The reference's smem_register kernels keep an xy tile in shared memory and
// each thread's z column in registers (src/stencil_kernel_smem_register.cu,
// src/prefetch_smem.cu:151-258).  Here:
//   * a workgroup of WY waves owns a frame of WY consecutive rows (one row
//     per wave, the whole row: each lane holds CX = 4*H cells as H quads of
//     4 consecutive columns 256 apart, so every 16-byte access of a wave
//     covers 1 KiB contiguously) and marches it along z over a chunk;
//   * per plane, every wave writes its row to an LDS plane, and waves
//     0..2R-1 also load and write one of the frame's 2R halo rows (R above,
//     R below), so y neighbours come from LDS with one barrier per plane and
//     no row is read twice by the workgroup; the LDS plane is double
//     buffered;
//   * x neighbours at distance d <= 4 lie in the lane's own quad or in the
//     adjacent lane's quad: wave-wide DPP rotates (`wave_ror:1`,
//     `wave_rol:1`), with lane 0 / lane 63 taking the neighbouring quad of
//     the other end of the wave;
//   * z: a ring of the R previous planes and R running partial sums.  When
//     plane m arrives its x and y terms and the ring (planes m-R..m-1) start
//     P_m in the reference's order (i-prev d=R..1, i-next 1..R, j-prev,
//     j-next, k-prev), and m is added to P_{m-1}..P_{m-R} as their next
//     k-next term; P_{m-R} is then complete.
// Level 0 of the next plane (own row and halo row) is prefetched into
// registers before the current plane is processed.  Only interior cells are
// stored; the output's shell is untouched.
//
// The same kernel serves the COOP variant (src/stencil_kernel_coop.cu:5-49:
// every sweep in one launch, a grid-wide barrier after each): launched
// cooperatively with `sweeps` > 1, each workgroup runs its tiles of a sweep,
// waits at the grid barrier and starts the next sweep with the buffers
// swapped.
//
// Two more modes serve the 16-bit storage scheme of mixed.hip (JACOBI, FP32
// arithmetic, T = float): the correction sweep reads 16-bit e, widens it,
// and writes round16(widen(d) + w*sum) (d read at the output cell), the
// defect reads FP32 x and writes round16(w*sum - x).
#include <hip/hip_cooperative_groups.h>

#include <type_traits>

#include "st_device.hpp"
#include "st_half.hpp"
#include "st_internal.hpp"

namespace st {

struct RowsArgs {
    int nx, ny, nz;
    long long sz;   // nx*ny
    int wx;         // waves per row (4 cells per lane when > 1)
    int wy;         // core rows per frame
    int tiles_y, zchunk, ntiles;
    int nt;         // non-temporal stores
    int sweeps;     // sweeps of this launch (> 1 only for cooperative launches)
    const int *eq;  // sweep mode: 1 when both buffers' shells are equal
                    // (edge quads then store whole, shell cells unchanged)
    // defect and fold modes of the consecutive layout (null: not written):
    // eout receives round16(widen(d) + 0), the first correction sweep from
    // a zero correction; xout the folded field x' (fold mode)
    uint16_t *eout;
    float *xout;
};

template <typename T> struct Quad;
template <> struct Quad<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Quad<double> { typedef double type __attribute__((ext_vector_type(4))); };

// Wave-wide rotates read a valid lane for every lane, so bound_ctrl (zero for
// an invalid source lane) never applies; setting it lets the compiler drop
// the copy of the old value.
template <typename T>
__device__ __forceinline__ T dpp_prev(T v)  // lane i <- lane i-1 (rotating)
{
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x13C, 0xF, 0xF, true));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x13C, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x13C, 0xF, 0xF, true);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

template <typename T>
__device__ __forceinline__ T dpp_next(T v)  // lane i <- lane i+1 (rotating)
{
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, true));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x134, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x134, 0xF, 0xF, true);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

// Kernel modes: a sweep, the two 16-bit modes, and RM_FOLD (consecutive
// layout only): x' = x + widen(e) on the fly, x' stored, then the defect of
// x' (the fold of a restart fused into its defect sweep).
enum : int { RM_SWEEP = 0, RM_CORR = 1, RM_DEFECT = 2, RM_FOLD = 3 };

// VGPRs a lane needs with the ring and partials shifted every plane (an
// upper bound for the phase-unrolled loop; shapes above 250 are not built),
// and the workgroup size whose register budget the kernel fits: 1024
// threads leave 128 VGPRs, 768 leave 168, 512 leave 256.

template <int R, int CX, typename T, int MODE = RM_SWEEP, int HS = 1>
constexpr int rows_vgprs()
{
    return (9 * R + 3) * CX * (int)(sizeof(T) / 4) / 2 + 25 + (MODE == RM_CORR ? CX : 0) +
           (HS - 1) * (CX * (int)(sizeof(T) / 4) + 8);
}

template <int R, int CX, typename T, int MODE = RM_SWEEP, int HS = 1>
constexpr bool rows_supported() { return rows_vgprs<R, CX, T, MODE, HS>() <= 250; }

// Workgroup size from the compiler's allocation of the phase-unrolled plane
// loop (no spills at these bounds): every FP32 sweep and the 4-cell 16-bit
// modes fit the 128 VGPRs of 16 waves (radius 4, 8 cells: 126); the 8-cell
// 16-bit modes at radius 4 take 12 waves, FP64 12 waves with 4 cells
// (radius 4: 148 VGPRs) or with 8 cells at radius 1, else 8.
template <int R, int CX, typename T, int MODE = RM_SWEEP, int HS = 1>
constexpr int rows_max_threads()
{
    if constexpr (sizeof(T) == 8) {
        return CX == 4 ? (R <= 3 ? 1024 : 768) : (R == 1 ? 768 : 512);
    } else if constexpr (MODE == RM_SWEEP) {
        return 1024;
    } else {
        return CX == 4 || R <= 3 ? 1024 : 768;
    }
}

template <int OP, int R, int CX, typename T, int MODE = RM_SWEEP, bool F16 = false, int HS = 1>
__global__ __launch_bounds__((rows_max_threads<R, CX, T, MODE, HS>())) void k_rows(
    const void *__restrict__ in0, void *__restrict__ out0, const uint16_t *__restrict__ aux, RowsArgs a,
    T kc)
{
    static_assert(R >= 1 && R <= 4, "x neighbours come from the adjacent lane's quad");
    static_assert(MODE == RM_SWEEP || sizeof(T) == 4, "the 16-bit modes compute in FP32");
    constexpr int H = CX / 4;
    typedef typename Quad<T>::type QV;
    typedef uint16_t U4 __attribute__((ext_vector_type(4)));
    // prefetch storage: 16-bit inputs as packed words
    typedef typename std::conditional<MODE == RM_CORR, uint32_t, T>::type P;
    constexpr int NP = MODE == RM_CORR ? CX / 2 : CX;
    // LDS quads: an FP64 quad is kept as two 16-byte halves in separate
    // 64-lane arrays, so every 16-byte access of a wave is contiguous (with
    // the 32-byte lane stride of whole quads each access was a two-way bank
    // conflict)
    constexpr int NH = (int)sizeof(QV) / 16;
    typedef T HQ __attribute__((ext_vector_type(4 / NH)));
    extern __shared__ __align__(16) unsigned char lds_raw[];
    HQ *lds = reinterpret_cast<HQ *>(lds_raw);
    auto lds_put = [](HQ *b, int qi, int l, const T *v) {
#pragma unroll
        for (int hh = 0; hh < NH; ++hh) {
            HQ x;
#pragma unroll
            for (int j = 0; j < 4 / NH; ++j) x[j] = v[hh * (4 / NH) + j];
            b[(qi * NH + hh) * 64 + l] = x;
        }
    };
    auto lds_get = [](const HQ *b, int qi, int l, T *v) {
#pragma unroll
        for (int hh = 0; hh < NH; ++hh) {
            const HQ x = b[(qi * NH + hh) * 64 + l];
#pragma unroll
            for (int j = 0; j < 4 / NH; ++j) v[hh * (4 / NH) + j] = x[j];
        }
    };
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // 8 cells per lane are laid out only for rows of one wave, so the seam
    // exchange of multi-wave rows compiles away there
    const int wx = CX == 8 ? 1 : a.wx, wy = a.wy;
    const int wr = w / wx, wxp = w - wr * wx;  // row of the frame, part of the row
    const int QW = wx * H;                     // quads per LDS row
    const int lrows = wy + 2 * R;              // LDS rows per buffer: R halo, wy core, R halo
    const bool first = lane == 0, last = lane == 63;
    const bool eqs = MODE == RM_SWEEP && a.eq != nullptr && *a.eq != 0;

    // own quads: columns 256*(wxp*H + h) + 4*lane
    int xo[H];
    bool act[H], intq[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int x = 256 * (wxp * H + h) + 4 * lane;
        act[h] = x < a.nx;
        xo[h] = act[h] ? x : a.nx - 4;  // quads past the row reload its last cells
        intq[h] = act[h] && x >= R && x + 4 <= a.nx - R;
    }
    // halo strips of this wave: strip s = w + i*waves (i < HS) of the 2R*wx
    // strips, row hr = s / wx of the 2R halo rows, part hp = s % wx of it
    const int waves = blockDim.x >> 6;
    bool has_halo[HS];
    int hr[HS], hp[HS], hxo[HS][H];
#pragma unroll
    for (int i = 0; i < HS; ++i) {
        const int st = w + i * waves;
        has_halo[i] = st < 2 * R * wx;
        hr[i] = st / wx;
        hp[i] = st - hr[i] * wx;
#pragma unroll
        for (int h = 0; h < H; ++h) hxo[i][h] = min(256 * (hp[i] * H + h) + 4 * lane, a.nx - 4);
    }
    const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
    const int nsweeps = MODE == RM_SWEEP ? a.sweeps : 1;

    for (int sw = 0; sw < nsweeps; ++sw) {
    // odd sweeps read the first output and write the first input
    const void *in_ = (sw & 1) ? static_cast<const void *>(out0) : in0;
    void *out_ = (sw & 1) ? const_cast<void *>(in0) : out0;
    if (sw > 0) cooperative_groups::this_grid().sync();
    for (int tile = (int)bid; tile < a.ntiles; tile += (int)gridDim.x) {
        const int tyi = tile % a.tiles_y, tzi = tile / a.tiles_y;
        const int fy0 = tyi * wy;
        const int gy = fy0 + wr;  // own (core) row
        const int own_off = min(gy, a.ny - 1) * a.nx;
        const bool store_row = gy >= R && gy < a.ny - R;
        // halo rows: frame rows -R..-1 (LDS rows 0..R-1), wy..wy+R-1 (LDS
        // rows R+wy..)
        int hrow[HS], halo_off[HS];
#pragma unroll
        for (int i = 0; i < HS; ++i) {
            hrow[i] = hr[i] < R ? hr[i] : wy + hr[i];
            const int hgy = hr[i] < R ? fy0 - R + hr[i] : fy0 + wy + (hr[i] - R);
            halo_off[i] = min(max(hgy, 0), a.ny - 1) * a.nx;
        }
        const int z0 = tzi * a.zchunk, z1 = min(z0 + a.zchunk, a.nz);

        // prefetched rows are held as loaded (16-bit quads as two packed
        // words, widened when the plane is consumed: widening at the load
        // made each step wait for its own prefetch)
        auto load_row = [&](int k, int off, const int (&cols)[H], P (&d)[NP]) {
            k = min(max(k, 0), a.nz - 1);
            if constexpr (MODE == RM_CORR) {
                const uint16_t *p = static_cast<const uint16_t *>(in_) + (long long)k * a.sz + off;
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const uint2 q = *reinterpret_cast<const uint2 *>(p + cols[h]);
                    d[2 * h] = q.x;
                    d[2 * h + 1] = q.y;
                }
            } else {
                const T *p = static_cast<const T *>(in_) + (long long)k * a.sz + off;
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const QV q = *reinterpret_cast<const QV *>(p + cols[h]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) d[4 * h + j] = q[j];
                }
            }
        };
        const int k_out_lo = max(z0, R), k_out_hi = min(z1, a.nz - R);

        auto widen_row = [&](const P (&src)[NP], T (&d)[CX]) {
            if constexpr (MODE == RM_CORR) {
#pragma unroll
                for (int v = 0; v < NP; ++v) {
                    d[2 * v] = wid16<F16>((uint16_t)(src[v] & 0xffffu));
                    d[2 * v + 1] = wid16<F16>((uint16_t)(src[v] >> 16));
                }
            } else {
#pragma unroll
                for (int c = 0; c < CX; ++c) d[c] = src[c];
            }
        };

        T ring[R][CX], part[R][CX];
        P pf[NP], ph[HS][NP];
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < CX; ++c) ring[i][c] = part[i][c] = (T)0;
        const int kin_lo = z0 - R, kin_hi = z1 + R;
        load_row(kin_lo, own_off, xo, pf);
#pragma unroll
        for (int i = 0; i < HS; ++i)
            if (has_halo[i]) load_row(kin_lo, halo_off[i], hxo[i], ph[i]);
        // correction sweep: the defect at the cells a step completes (plane
        // m-R), loaded one step ahead, packed like the prefetched rows
        uint2 dqn[H];
        auto load_defect = [&](int k) {
            if constexpr (MODE == RM_CORR) {
                if (store_row && k >= k_out_lo && k < k_out_hi) {
                    const uint16_t *p = aux + (long long)k * a.sz + own_off;
#pragma unroll
                    for (int h = 0; h < H; ++h) dqn[h] = *reinterpret_cast<const uint2 *>(p + xo[h]);
                }
            }
        };
        load_defect(kin_lo - R);

        // planes in phases of R: the ring and the partial sums rotate
        // through their R slots (logical entry i of phase rp in slot
        // (rp + i) % R), so no values are shifted; the scheduling barrier
        // keeps each plane's work in place, which also keeps the register
        // need at the one-plane level
        for (int m0 = kin_lo; m0 < kin_hi; m0 += R) {
#pragma unroll
        for (int rp = 0; rp < R; ++rp) {
            const int m = m0 + rp;
            if (m >= kin_hi) break;
            __builtin_amdgcn_sched_barrier(0);
            T nw[CX];
            widen_row(pf, nw);
            HQ *buf = lds + (size_t)((m & 1) * lrows) * QW * NH * 64;
#pragma unroll
            for (int h = 0; h < H; ++h) lds_put(buf, (R + wr) * QW + wxp * H + h, lane, nw + 4 * h);
#pragma unroll
            for (int i = 0; i < HS; ++i) {
                if (!has_halo[i]) continue;
                T hv[CX];
                widen_row(ph[i], hv);
#pragma unroll
                for (int h = 0; h < H; ++h) lds_put(buf, hrow[i] * QW + hp[i] * H + h, lane, hv + 4 * h);
            }
            const int k = m - R;
            const bool out_plane = store_row && k >= k_out_lo && k < k_out_hi;
            uint2 dq[H];
            if constexpr (MODE == RM_CORR) {
#pragma unroll
                for (int h = 0; h < H; ++h) dq[h] = dqn[h];
            }
            if (m + 1 < kin_hi) {
                load_row(m + 1, own_off, xo, pf);
#pragma unroll
                for (int i = 0; i < HS; ++i)
                    if (has_halo[i]) load_row(m + 1, halo_off[i], hxo[i], ph[i]);
                load_defect(k + 1);
            }
            __syncthreads();

            // x terms of plane m: own quad, else the adjacent lane's quad
            T pq[H][4], nq[H][4];
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    pq[h][j] = dpp_prev(nw[4 * h + j]);
                    nq[h][j] = dpp_next(nw[4 * h + j]);
                }
            // lane 0 of the first quad and lane 63 of the last take the
            // neighbouring wave's edge quad when the row spans several waves
            // (LDS broadcast reads), else the rotate's wrap (unused values
            // next to the shell)
            T lq[4], rq[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                lq[j] = pq[0][j];
                rq[j] = nq[H - 1][j];
            }
            if (wx > 1) {
                T le[4], re[4];
                lds_get(buf, (R + wr) * QW + max(wxp * H - 1, 0), 63, le);
                lds_get(buf, (R + wr) * QW + min(wxp * H + H, QW - 1), 0, re);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const T l0 = le[j], r0 = re[j];
                    lq[j] = wxp > 0 ? l0 : lq[j];
                    rq[j] = wxp < wx - 1 ? r0 : rq[j];
                }
            }
            T s[CX];
#pragma unroll
            for (int h = 0; h < H; ++h) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    T u = (T)0;
#pragma unroll
                    for (int d = R; d >= 1; --d) {
                        T v;
                        if (j - d >= 0) {
                            v = nw[4 * h + j - d];
                        } else {
                            const T mine = pq[h][4 + j - d];
                            const T other = h > 0 ? pq[h - 1][4 + j - d] : lq[4 + j - d];
                            v = first ? other : mine;
                        }
                        u += v;
                    }
#pragma unroll
                    for (int d = 1; d <= R; ++d) {
                        T v;
                        if (j + d < 4) {
                            v = nw[4 * h + j + d];
                        } else {
                            const T mine = nq[h][j + d - 4];
                            const T other = h < H - 1 ? nq[h + 1][j + d - 4] : rq[j + d - 4];
                            v = last ? other : mine;
                        }
                        u += v;
                    }
                    s[4 * h + j] = u;
                }
            }
            // y terms from the LDS plane: rows R+wr-d (d = R..1), then R+wr+d
#pragma unroll
            for (int d = R; d >= 1; --d) {
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    T q[4];
                    lds_get(buf, (R + wr - d) * QW + wxp * H + h, lane, q);
#pragma unroll
                    for (int j = 0; j < 4; ++j) s[4 * h + j] += q[j];
                }
            }
#pragma unroll
            for (int d = 1; d <= R; ++d) {
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    T q[4];
                    lds_get(buf, (R + wr + d) * QW + wxp * H + h, lane, q);
#pragma unroll
                    for (int j = 0; j < 4; ++j) s[4 * h + j] += q[j];
                }
            }
            // k-prev terms: logical ring[0] = plane m-R ... ring[R-1] = plane m-1
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CX; ++c) s[c] += ring[(rp + i) % R][c];
            // plane m is the next k-next term of P_{m-R} .. P_{m-1}
            // (logical part[0] = P_{m-R}, which it completes)
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CX; ++c) part[(rp + i) % R][c] += nw[c];
            // P_{m-R} is complete: plane k = m-R, centre logical ring[0]
            if (out_plane) {
                if constexpr (MODE == RM_SWEEP) {
                    T res[CX];
#pragma unroll
                    for (int c = 0; c < CX; ++c) res[c] = finish<OP>(part[rp][c], ring[rp][c], kc);
                    T *row = static_cast<T *>(out_) + (long long)k * a.sz + (long long)gy * a.nx;
#pragma unroll
                    for (int h = 0; h < H; ++h) {
                        if (intq[h]) {
                            QV q;
#pragma unroll
                            for (int j = 0; j < 4; ++j) q[j] = res[4 * h + j];
                            if (a.nt) __builtin_nontemporal_store(q, reinterpret_cast<QV *>(row + xo[h]));
                            else *reinterpret_cast<QV *>(row + xo[h]) = q;
                        } else if (act[h] && eqs) {
                            // equal shells: the whole quad, shell cells with
                            // the value they hold (full cache lines)
                            QV q;
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                q[j] = interior1(xo[h] + j, a.nx, R) ? res[4 * h + j] : ring[rp][4 * h + j];
                            if (a.nt) __builtin_nontemporal_store(q, reinterpret_cast<QV *>(row + xo[h]));
                            else *reinterpret_cast<QV *>(row + xo[h]) = q;
                        } else if (act[h]) {
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (interior1(xo[h] + j, a.nx, R)) row[xo[h] + j] = res[4 * h + j];
                        }
                    }
                } else {
                    // 16-bit output: correction round16(widen(d) + w*sum),
                    // defect round16(w*sum - x)
                    U4 res[H];
#pragma unroll
                    for (int h = 0; h < H; ++h)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const float sw = part[rp][4 * h + j] * kc;
                            if constexpr (MODE == RM_CORR) {
                                const uint32_t wd = j < 2 ? dq[h].x : dq[h].y;
                                const uint16_t dc = (uint16_t)(j & 1 ? wd >> 16 : wd & 0xffffu);
                                res[h][j] = rnd16<F16>(wid16<F16>(dc) + sw);
                            }
                            else res[h][j] = rnd16<F16>(sw - ring[rp][4 * h + j]);
                        }
                    uint16_t *row = static_cast<uint16_t *>(out_) + (long long)k * a.sz + (long long)gy * a.nx;
#pragma unroll
                    for (int h = 0; h < H; ++h) {
                        if (intq[h]) {
                            *reinterpret_cast<U4 *>(row + xo[h]) = res[h];
                        } else if (act[h]) {
                            // the 16-bit outputs' shells are zero (the
                            // mixed driver fills them and never writes them),
                            // so edge quads store whole with zero there
                            U4 q;
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                q[j] = interior1(xo[h] + j, a.nx, R) ? res[h][j] : (uint16_t)0;
                            *reinterpret_cast<U4 *>(row + xo[h]) = q;
                        }
                    }
                }
            }
            // plane m and P_m take the slot of plane m-R and P_{m-R}
#pragma unroll
            for (int c = 0; c < CX; ++c) {
                ring[rp][c] = nw[c];
                part[rp][c] = s[c];
            }
        }
        }
        __syncthreads();  // the next tile rewrites the LDS planes
    }
    }
}

// ---------------------------------------------------------------- consecutive layout
//
// Rows of nx = 64*CX columns (CX = 4, 6 or 8; 256, 384 or 512 columns) are
// held by one wave with CX consecutive columns per lane, so every lane is
// busy (the quad layout above leaves a quarter of the lanes idle at 384
// columns).  Loads, stores and LDS accesses are vectors of V consecutive
// values; the x neighbours at distance d <= R <= CX are the lane's own
// values or the adjacent lane's, one DPP rotate per value (R per side); the
// rotate's wrap at lanes 0 and 63 reaches only cells of the x shell, which
// are never computed.  Everything else (LDS plane per step, halo strips, z
// ring and partial sums in phases of R, the 16-bit modes) is as in k_rows.
template <typename E, int V> struct VecN { typedef E type __attribute__((ext_vector_type(V))); };

// values per vector access: the largest power of two dividing CX with at
// most 16 bytes
template <int CX, int ES>
constexpr int rowc_vec() { return (CX % 8 == 0 && 8 * ES <= 16) ? 8 : (CX % 4 == 0 && 4 * ES <= 16) ? 4 : 2; }

// VGPRs the lane needs, in 32-bit words: ring and partial sums 2R*CX
// values, the plane, its prefetch, the halo prefetch, the sums, the x
// neighbours and the vector temporaries about 6*CX, and 48 words of
// addresses and loop state (fitted to the compiler's allocation of every
// instantiation).  The workgroup size follows from it: 4, 3 or 2 waves per
// SIMD (128, 168 or 256 VGPRs, kept with a margin); beyond 200 words the
// kernel is not built and the quad layout serves.
template <int R, int CX, typename T>
constexpr int rowc_words() { return (2 * R + 6) * CX * (int)(sizeof(T) / 4) + 48; }

template <int R, int CX, typename T>
constexpr int rowc_max_threads()
{
    constexpr int w = rowc_words<R, CX, T>();
    return w <= 128 ? 1024 : w <= 150 ? 768 : w <= 200 ? 512 : 0;
}

// N consecutive 32-bit words from a 4-byte aligned address, in vectors of
// 4, 2 or 1 words.
template <int N>
__device__ __forceinline__ void load_words(const uint32_t *__restrict__ p, uint32_t (&d)[N])
{
    constexpr int V = N % 4 == 0 ? 4 : N % 2 == 0 ? 2 : 1;
#pragma unroll
    for (int v = 0; v < N / V; ++v) {
        if constexpr (V == 1) {
            d[v] = p[v];
        } else {
            typedef typename VecN<uint32_t, V>::type W;
            const W q = *reinterpret_cast<const W *>(p + v * V);
#pragma unroll
            for (int j = 0; j < V; ++j) d[v * V + j] = q[j];
        }
    }
}

template <int OP, int R, int CX, typename T, int MODE, bool F16>
__global__ __launch_bounds__((rowc_max_threads<R, CX, T>())) void k_rowc(
    const void *__restrict__ in0, void *__restrict__ out0, const uint16_t *__restrict__ aux, RowsArgs a, T kc)
{
    static_assert(R >= 1 && R <= CX, "x neighbours come from the adjacent lane");
    static_assert(MODE == RM_SWEEP || sizeof(T) == 4, "the 16-bit modes compute in FP32");
    constexpr bool FOLD = MODE == RM_FOLD;
    constexpr int NE = FOLD ? CX / 2 : 1;  // packed correction words per row (fold)
    // storage type of the input field
    typedef typename std::conditional<MODE == RM_CORR, uint16_t, T>::type S;
    // prefetch storage: 16-bit inputs as packed words
    typedef typename std::conditional<MODE == RM_CORR, uint32_t, T>::type P;
    constexpr int NP = MODE == RM_CORR ? CX / 2 : CX;
    constexpr int VS = rowc_vec<CX, (int)sizeof(S)>();  // global loads
    constexpr int VT = rowc_vec<CX, (int)sizeof(T)>();  // LDS and sweep stores
    constexpr int VH = rowc_vec<CX, 2>();                // 16-bit stores and defect loads
    typedef typename VecN<S, VS>::type SV;
    typedef typename VecN<T, VT>::type TV;
    typedef typename VecN<uint16_t, VH>::type HV;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wy = a.wy;
    constexpr int RS = 64 * CX;   // LDS row stride (values)
    const int lrows = wy + 2 * R;  // LDS rows per buffer: R halo, wy core, R halo
    const int xo = lane * CX;      // first column of the lane
    const bool act = xo < a.nx;
    const int xl = act ? xo : 0;   // load column (idle lanes read the row start)
    const bool whole = xo >= R && xo + CX <= a.nx - R;  // every cell interior
    const bool eqs = MODE == RM_SWEEP && a.eq != nullptr && *a.eq != 0;
    // halo strip of this wave: waves 0..2R-1 each load one of the frame's
    // 2R halo rows (rows -R..-1 into LDS rows 0..R-1, wy..wy+R-1 into
    // R+wy..)
    const bool has_halo = w < 2 * R;
    const int hrow = w < R ? w : wy + w;
    const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
    const int nsweeps = MODE == RM_SWEEP ? a.sweeps : 1;

    for (int sw = 0; sw < nsweeps; ++sw) {
    const void *in_ = (sw & 1) ? static_cast<const void *>(out0) : in0;
    void *out_ = (sw & 1) ? const_cast<void *>(in0) : out0;
    if (sw > 0) cooperative_groups::this_grid().sync();
    for (int tile = (int)bid; tile < a.ntiles; tile += (int)gridDim.x) {
        const int tyi = tile % a.tiles_y, tzi = tile / a.tiles_y;
        const int fy0 = tyi * wy;
        const int gy = fy0 + w;
        const int own_off = min(gy, a.ny - 1) * a.nx;
        const bool store_row = gy >= R && gy < a.ny - R;
        const int hgy = w < R ? fy0 - R + w : fy0 + wy + (w - R);
        const int halo_off = min(max(hgy, 0), a.ny - 1) * a.nx;
        const int z0 = tzi * a.zchunk, z1 = min(z0 + a.zchunk, a.nz);
        const int k_out_lo = max(z0, R), k_out_hi = min(z1, a.nz - R);

        // Prefetched rows are held as loaded: 16-bit values stay packed two
        // per word and are widened when the plane is consumed (widened at
        // the load, every conversion waited for the load in flight, which
        // made each step wait for its own prefetch)
        auto load_row = [&](int k, int off, P (&d)[NP]) {
            k = min(max(k, 0), a.nz - 1);
            if constexpr (MODE == RM_CORR) {
                const uint32_t *p = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(in_) +
                                                                       (long long)k * a.sz + off + xl);
                load_words<NP>(p, d);
            } else {
                const S *p = static_cast<const S *>(in_) + (long long)k * a.sz + off + xl;
#pragma unroll
                for (int v = 0; v < CX / VS; ++v) {
                    const SV q = *reinterpret_cast<const SV *>(p + v * VS);
#pragma unroll
                    for (int j = 0; j < VS; ++j) d[v * VS + j] = q[j];
                }
            }
        };
        auto load_e = [&](int k, int off, uint32_t (&d)[NE]) {
            if constexpr (FOLD) {
                k = min(max(k, 0), a.nz - 1);
                load_words<NE>(reinterpret_cast<const uint32_t *>(aux + (long long)k * a.sz + off + xl), d);
            }
        };
        // x + widen(e), in the order of the separate fold (fold mode)
        auto fold_row = [&](T (&d)[CX], const uint32_t (&e)[NE]) {
            if constexpr (FOLD) {
#pragma unroll
                for (int c = 0; c < CX; ++c)
                    d[c] = d[c] + wid16<F16>((uint16_t)(c & 1 ? e[c / 2] >> 16 : e[c / 2] & 0xffffu));
            }
        };
        auto widen_row = [&](const P (&src)[NP], T (&d)[CX]) {
            if constexpr (MODE == RM_CORR) {
#pragma unroll
                for (int v = 0; v < NP; ++v) {
                    d[2 * v] = wid16<F16>((uint16_t)(src[v] & 0xffffu));
                    d[2 * v + 1] = wid16<F16>((uint16_t)(src[v] >> 16));
                }
            } else {
#pragma unroll
                for (int c = 0; c < CX; ++c) d[c] = src[c];
            }
        };

        T ring[R][CX], part[R][CX];
        P pf[NP], ph[NP];
        uint32_t pe[NE], phe[NE];
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < CX; ++c) ring[i][c] = part[i][c] = (T)0;
        const int kin_lo = z0 - R, kin_hi = z1 + R;
        load_row(kin_lo, own_off, pf);
        load_e(kin_lo, own_off, pe);
        if (has_halo) {
            load_row(kin_lo, halo_off, ph);
            load_e(kin_lo, halo_off, phe);
        }
        // correction sweep: the defect of the cells a step completes (plane
        // m-R), loaded one step ahead, packed like the prefetched rows
        uint32_t dqn[CX / 2];
        auto load_defect = [&](int k) {
            if constexpr (MODE == RM_CORR) {
                if (store_row && k >= k_out_lo && k < k_out_hi)
                    load_words<CX / 2>(reinterpret_cast<const uint32_t *>(aux + (long long)k * a.sz + own_off + xl),
                                       dqn);
            }
        };
        load_defect(kin_lo - R);

        for (int m0 = kin_lo; m0 < kin_hi; m0 += R) {
#pragma unroll
        for (int rp = 0; rp < R; ++rp) {
            const int m = m0 + rp;
            if (m >= kin_hi) break;
            __builtin_amdgcn_sched_barrier(0);
            T nw[CX];
            widen_row(pf, nw);
            fold_row(nw, pe);
            T *buf = lds + (size_t)((m & 1) * lrows) * RS;
#pragma unroll
            for (int v = 0; v < CX / VT; ++v) {
                TV q;
#pragma unroll
                for (int j = 0; j < VT; ++j) q[j] = nw[v * VT + j];
                *reinterpret_cast<TV *>(buf + (R + w) * RS + xo + v * VT) = q;
            }
            if (has_halo) {
                T hv[CX];
                widen_row(ph, hv);
                fold_row(hv, phe);
#pragma unroll
                for (int v = 0; v < CX / VT; ++v) {
                    TV q;
#pragma unroll
                    for (int j = 0; j < VT; ++j) q[j] = hv[v * VT + j];
                    *reinterpret_cast<TV *>(buf + hrow * RS + xo + v * VT) = q;
                }
            }
            const int k = m - R;
            const bool out_plane = store_row && k >= k_out_lo && k < k_out_hi;
            uint32_t dq[CX / 2];
            if constexpr (MODE == RM_CORR) {
#pragma unroll
                for (int v = 0; v < CX / 2; ++v) dq[v] = dqn[v];
            }
            if (m + 1 < kin_hi) {
                load_row(m + 1, own_off, pf);
                load_e(m + 1, own_off, pe);
                if (has_halo) {
                    load_row(m + 1, halo_off, ph);
                    load_e(m + 1, halo_off, phe);
                }
                load_defect(k + 1);
            }
            __syncthreads();

            // x terms: own values, else the adjacent lane's (pv[i] holds
            // column xo-R+i, nv[i] column xo+CX+i)
            T pv[R], nv[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                pv[i] = dpp_prev(nw[CX - R + i]);
                nv[i] = dpp_next(nw[i]);
            }
            T s[CX];
#pragma unroll
            for (int c = 0; c < CX; ++c) {
                T u = (T)0;
#pragma unroll
                for (int d = R; d >= 1; --d) u += c - d >= 0 ? nw[c - d] : pv[R + c - d];
#pragma unroll
                for (int d = 1; d <= R; ++d) u += c + d < CX ? nw[c + d] : nv[c + d - CX];
                s[c] = u;
            }
            // y terms from the LDS plane: rows R+w-d (d = R..1), then R+w+d
#pragma unroll
            for (int d = R; d >= 1; --d) {
#pragma unroll
                for (int v = 0; v < CX / VT; ++v) {
                    const TV q = *reinterpret_cast<const TV *>(buf + (R + w - d) * RS + xo + v * VT);
#pragma unroll
                    for (int j = 0; j < VT; ++j) s[v * VT + j] += q[j];
                }
            }
#pragma unroll
            for (int d = 1; d <= R; ++d) {
#pragma unroll
                for (int v = 0; v < CX / VT; ++v) {
                    const TV q = *reinterpret_cast<const TV *>(buf + (R + w + d) * RS + xo + v * VT);
#pragma unroll
                    for (int j = 0; j < VT; ++j) s[v * VT + j] += q[j];
                }
            }
            // k-prev terms: logical ring[i] = plane m-R+i, in slot (rp+i) % R
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CX; ++c) s[c] += ring[(rp + i) % R][c];
            // plane m is the next k-next term of P_{m-R} .. P_{m-1}
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CX; ++c) part[(rp + i) % R][c] += nw[c];
            // P_{m-R} (slot rp) is complete: plane k, centre ring slot rp
            if (out_plane && act) {
                if constexpr (MODE == RM_SWEEP) {
                    T res[CX];
#pragma unroll
                    for (int c = 0; c < CX; ++c) res[c] = finish<OP>(part[rp][c], ring[rp][c], kc);
                    T *row = static_cast<T *>(out_) + (long long)k * a.sz + (long long)gy * a.nx + xo;
                    if (whole || eqs) {
                        // equal shells: shell cells with the value they
                        // hold, so rows go out as whole lines
#pragma unroll
                        for (int v = 0; v < CX / VT; ++v) {
                            TV q;
#pragma unroll
                            for (int j = 0; j < VT; ++j) {
                                const int c = v * VT + j;
                                q[j] = whole || interior1(xo + c, a.nx, R) ? res[c] : ring[rp][c];
                            }
                            *reinterpret_cast<TV *>(row + v * VT) = q;
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < CX; ++c)
                            if (interior1(xo + c, a.nx, R)) row[c] = res[c];
                    }
                } else {
                    // 16-bit output: correction round16(widen(d) + w*sum),
                    // defect round16(w*sum - x); the shells of the 16-bit
                    // outputs are zero, so whole vectors go out
                    uint16_t res[CX];
#pragma unroll
                    for (int c = 0; c < CX; ++c) {
                        const float sw = part[rp][c] * kc;
                        uint16_t r16;
                        if constexpr (MODE == RM_CORR) {
                            const uint16_t dc = (uint16_t)(c & 1 ? dq[c / 2] >> 16 : dq[c / 2] & 0xffffu);
                            r16 = rnd16<F16>(wid16<F16>(dc) + sw);
                        }
                        else r16 = rnd16<F16>(sw - ring[rp][c]);
                        res[c] = whole || interior1(xo + c, a.nx, R) ? r16 : (uint16_t)0;
                    }
                    const long long ro = (long long)k * a.sz + (long long)gy * a.nx + xo;
                    auto store16 = [&](uint16_t *row, const uint16_t (&v16)[CX]) {
#pragma unroll
                        for (int v = 0; v < CX / VH; ++v) {
                            HV q;
#pragma unroll
                            for (int j = 0; j < VH; ++j) q[j] = v16[v * VH + j];
                            *reinterpret_cast<HV *>(row + v * VH) = q;
                        }
                    };
                    store16(static_cast<uint16_t *>(out_) + ro, res);
                    if constexpr (MODE != RM_CORR) {
                        if (a.eout) {
                            // the first correction sweep from a zero
                            // correction: round16(widen(d) + w*0), shell 0
                            uint16_t e1[CX];
#pragma unroll
                            for (int c = 0; c < CX; ++c)
                                e1[c] = whole || interior1(xo + c, a.nx, R) ? rnd16<F16>(wid16<F16>(res[c]) + 0.0f)
                                                                           : (uint16_t)0;
                            store16(a.eout + ro, e1);
                        }
                    }
                    if constexpr (FOLD) {
                        // x' of the row, whole (shell cells: x + widen(0))
#pragma unroll
                        for (int v = 0; v < CX / VT; ++v) {
                            TV q;
#pragma unroll
                            for (int j = 0; j < VT; ++j) q[j] = ring[rp][v * VT + j];
                            *reinterpret_cast<TV *>(a.xout + ro + v * VT) = q;
                        }
                    }
                }
            }
            // plane m and P_m take the slot of plane m-R and P_{m-R}
#pragma unroll
            for (int c = 0; c < CX; ++c) {
                ring[rp][c] = nw[c];
                part[rp][c] = s[c];
            }
        }
        }
        __syncthreads();  // the next tile rewrites the LDS planes
    }
    }
}

// ---------------------------------------------------------------- host side

// Built instantiations: every (R, CX, T) of the sweep whose registers fit,
// the 16-bit modes likewise, and two halo strips per wave (frames shorter
// than 2R rows: rows spanning several waves) for the 4-cell sweeps.
template <int R, int CX, typename T, int MODE, bool F16, int HS>
static const void *rows_fn_t(int op)
{
    if constexpr (!rows_supported<R, CX, T, MODE, HS>() || (HS > 1 && (CX != 4 || MODE != RM_SWEEP))) {
        return nullptr;
    } else if constexpr (MODE != RM_SWEEP) {
        return reinterpret_cast<const void *>(&k_rows<OP_JACOBI, R, CX, T, MODE, F16, HS>);
    } else {
        return op == OP_THESIS ? reinterpret_cast<const void *>(&k_rows<OP_THESIS, R, CX, T, MODE, F16, HS>)
                               : reinterpret_cast<const void *>(&k_rows<OP_JACOBI, R, CX, T, MODE, F16, HS>);
    }
}

template <int R, typename T, int MODE, bool F16>
static const void *rows_fn_r(int op, int cx, int hs, int *threads)
{
    if (cx == 8) {
        *threads = rows_max_threads<R, 8, T, MODE>();
        return rows_fn_t<R, 8, T, MODE, F16, 1>(op);
    }
    if (hs == 2) {
        *threads = rows_max_threads<R, 4, T, MODE, 2>();
        return rows_fn_t<R, 4, T, MODE, F16, 2>(op);
    }
    *threads = rows_max_threads<R, 4, T, MODE>();
    return rows_fn_t<R, 4, T, MODE, F16, 1>(op);
}

template <typename T, int MODE, bool F16>
static const void *rows_fn_d(int op, int r, int cx, int hs, int *threads)
{
    switch (r) {
    case 1: return rows_fn_r<1, T, MODE, F16>(op, cx, hs, threads);
    case 2: return rows_fn_r<2, T, MODE, F16>(op, cx, hs, threads);
    case 3: return rows_fn_r<3, T, MODE, F16>(op, cx, hs, threads);
    default: return rows_fn_r<4, T, MODE, F16>(op, cx, hs, threads);
    }
}

// mode RM_SWEEP: c.dtype F32/F64; RM_CORR / RM_DEFECT: FP32 arithmetic on
// 16-bit storage (f16 selects FP16, else BF16); hs halo strips per wave
static const void *rows_fn_mode(const Cfg &c, int cx, int mode, bool f16, int *threads, int hs = 1)
{
    *threads = 0;
    if (c.r < 1 || c.r > 4) return nullptr;
    if (mode == RM_CORR)
        return f16 ? rows_fn_d<float, RM_CORR, true>(c.op, c.r, cx, hs, threads)
                   : rows_fn_d<float, RM_CORR, false>(c.op, c.r, cx, hs, threads);
    if (mode == RM_DEFECT)
        return f16 ? rows_fn_d<float, RM_DEFECT, true>(c.op, c.r, cx, hs, threads)
                   : rows_fn_d<float, RM_DEFECT, false>(c.op, c.r, cx, hs, threads);
    if (c.dtype == ST_F64) return rows_fn_d<double, RM_SWEEP, false>(c.op, c.r, cx, hs, threads);
    return rows_fn_d<float, RM_SWEEP, false>(c.op, c.r, cx, hs, threads);
}

// Consecutive layout: rows of exactly 64*CX columns, CX = 4, 6 or 8 and at
// least the radius.
static int rowc_cx(const Cfg &c)
{
    if (c.nx % 64 || c.r < 1 || c.r > 4) return 0;
    const int cx = c.nx / 64;
    return (cx == 4 || cx == 6 || cx == 8) && c.r <= cx ? cx : 0;
}

template <int R, int CX, typename T, int MODE, bool F16>
static const void *rowc_fn_t(int op, int *threads)
{
    *threads = rowc_max_threads<R, CX, T>();
    if constexpr (rowc_max_threads<R, CX, T>() == 0)
        return nullptr;
    else if constexpr (MODE != RM_SWEEP)
        return reinterpret_cast<const void *>(&k_rowc<OP_JACOBI, R, CX, T, MODE, F16>);
    else
        return op == OP_THESIS ? reinterpret_cast<const void *>(&k_rowc<OP_THESIS, R, CX, T, MODE, F16>)
                               : reinterpret_cast<const void *>(&k_rowc<OP_JACOBI, R, CX, T, MODE, F16>);
}

template <typename T, int MODE, bool F16>
static const void *rowc_fn_d(int op, int r, int cx, int *threads)
{
#define RC_CASE(RR, CC) \
    if (r == RR && cx == CC) return rowc_fn_t<RR, CC, T, MODE, F16>(op, threads);
    RC_CASE(1, 4) RC_CASE(2, 4) RC_CASE(3, 4) RC_CASE(4, 4)
    RC_CASE(1, 6) RC_CASE(2, 6) RC_CASE(3, 6) RC_CASE(4, 6)
    RC_CASE(1, 8) RC_CASE(2, 8) RC_CASE(3, 8) RC_CASE(4, 8)
#undef RC_CASE
    return nullptr;
}

static const void *rowc_fn_mode(const Cfg &c, int cx, int mode, bool f16, int *threads)
{
    *threads = 0;
    if (mode == RM_CORR)
        return f16 ? rowc_fn_d<float, RM_CORR, true>(c.op, c.r, cx, threads)
                   : rowc_fn_d<float, RM_CORR, false>(c.op, c.r, cx, threads);
    if (mode == RM_DEFECT)
        return f16 ? rowc_fn_d<float, RM_DEFECT, true>(c.op, c.r, cx, threads)
                   : rowc_fn_d<float, RM_DEFECT, false>(c.op, c.r, cx, threads);
    if (mode == RM_FOLD)
        return f16 ? rowc_fn_d<float, RM_FOLD, true>(c.op, c.r, cx, threads)
                   : rowc_fn_d<float, RM_FOLD, false>(c.op, c.r, cx, threads);
    if (c.dtype == ST_F64) return rowc_fn_d<double, RM_SWEEP, false>(c.op, c.r, cx, threads);
    return rowc_fn_d<float, RM_SWEEP, false>(c.op, c.r, cx, threads);
}

// The row kernel a resolved configuration launches.
static const void *rows_kernel(const Cfg &c, int mode, bool f16, int *threads)
{
    if (c.rowc) return rowc_fn_mode(c, c.cx, mode, f16, threads);
    return rows_fn_mode(c, c.cx, mode, f16, threads, c.hs);
}

// Lane layout: 4 cells per lane and one wave per row up to 256 columns; up
// to 512, 8 cells per lane for sweeps where that kernel is built (measured
// on 384^3 and 512^3: equal or faster at every radius, up to 1.19x at R=3)
// and ceil(nx/256) waves of 4 cells for the 16-bit modes (1.12x / 1.17x
// for BF16 / FP16 at 384^3 R=4); beyond 512, ceil(nx/256) waves of 4 cells.
// Returns false when no built kernel fits (a frame must hold the 2R halo
// strips: rows >= 2R).
static bool rows_layout(const Cfg &c, int mode, bool f16, int *cx, int *wx, int *threads,
                        const void **fn, int *hs)
{
    *hs = 1;
    if (c.nx % 4 || c.nx < 8 || c.r < 1 || c.r > 4) return false;
    int t8 = 0;
    const void *f8 = c.nx <= 512 ? rows_fn_mode(c, 8, mode, f16, &t8) : nullptr;
    // coarsen_x 4 or 8 with HEURISTIC=0 picks the layout explicitly
    const int want = (!c.heuristic && (c.cx == 4 || c.cx == 8)) ? c.cx : 0;
    if (c.nx <= 256 && want != 8) {
        *cx = 4;
        *wx = 1;
    } else if (f8 && (want == 8 || (want == 0 && mode == RM_SWEEP))) {
        *cx = 8;
        *wx = 1;
    } else {
        if (want == 8) return false;
        *cx = 4;
        *wx = (c.nx + 255) / 256;
    }
    *fn = *cx == 8 ? f8 : rows_fn_mode(c, 4, mode, f16, threads);
    if (*cx == 8) *threads = t8;
    if (!*fn) return false;
    if ((*threads / 64) / *wx >= 2 * c.r) return true;
    if (*cx != 4) return false;
    // frames shorter than 2R rows: each wave loads two halo strips
    *fn = rows_fn_mode(c, 4, mode, f16, threads, 2);
    *hs = 2;
    return *fn && (*threads / 64) / *wx >= 1 && 2 * (*threads / 64) >= 2 * c.r * *wx;
}

bool rows_eligible(const Cfg &c)
{
    if (!((c.dtype == ST_F32 || c.dtype == ST_F64) && c.dims == 3)) return false;
    int cx, wx, threads, hs;
    const void *fn;
    return rows_layout(c, RM_SWEEP, false, &cx, &wx, &threads, &fn, &hs);
}

static st_status rows_resolve_mode(Cfg &c, int mode, bool f16)
{
    int cx = 0, wx = 1, threads = 0, hs = 1;
    const void *fn = nullptr;
    // Rows of 256, 384 or 512 columns take the consecutive layout (the
    // library's choice, or coarsen_x = nx/64 with HEURISTIC=0): every lane
    // busy, one DPP rotate per neighbour value, no selects.
    const int ccx = c.dims == 3 ? rowc_cx(c) : 0;
    const bool consec = ccx && (c.heuristic || c.cx == ccx);
    c.rowc = 0;
    if (consec) {
        cx = ccx;
        fn = rowc_fn_mode(c, cx, mode, f16, &threads);
        if (fn && threads / 64 >= 2 * c.r) c.rowc = 1;
    }
    if (!c.rowc && mode == RM_FOLD)
        return fail(ST_EUNSUPPORTED, "the fused fold and defect sweep needs rows of 256, 384 or 512 columns");
    if (!c.rowc && (c.dims != 3 || !rows_layout(c, mode, f16, &cx, &wx, &threads, &fn, &hs)))
        return fail(ST_EUNSUPPORTED, "row-streaming form needs a 3D F32/F64 grid, radius 1..4 and nx a "
                                     "multiple of 4 whose rows fit a workgroup with its halo strips");
    const int rows_max = threads / 64 / wx;
    // every one of the 2R*wx halo strips needs a wave (hs strips per wave)
    const int rows_min = (2 * c.r * wx + hs * wx - 1) / (hs * wx);
    int wy = rows_max;
    if (!c.heuristic && c.by > 0) {
        // block_y = rows per frame
        if (c.by < rows_min || c.by > rows_max)
            return fail(ST_EINVAL, "row-streaming form: block_y (rows per frame) must be %d..%d", rows_min, rows_max);
        wy = c.by;
    } else if (wy > c.ny) {
        wy = c.ny < rows_min ? rows_min : c.ny;
    }
    const int es = (c.dtype == ST_F64 && mode == RM_SWEEP) ? 8 : 4;
    // LDS: two planes of (rows + 2R) rows of wx*64 lanes' cells
    auto lds_of = [&](int rows) { return 2 * (rows + 2 * c.r) * wx * cx * 64 * es; };
    // the heuristic frame shrinks until its double-buffered planes fit
    while ((c.heuristic || c.by <= 0) && lds_of(wy) > kMaxLds && wy > rows_min) --wy;
    c.lds = lds_of(wy);
    if (c.lds > kMaxLds) return fail(ST_ELDS, "row-streaming form needs %d bytes of LDS", c.lds);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const int tiles_y = (c.ny + wy - 1) / wy;
    int chunk = (!c.heuristic && c.bz > 0) ? c.bz : 0;
    if (chunk <= 0) {
        // The z chunk that minimises the busiest CU's work: rounds of tiles
        // (one workgroup per CU) x planes marched per tile (chunk + 2R).
        // Measured at 384^3 R=4: FP32 (48 frames) 5 chunks of 77, one
        // round, 2.54 ms per 20 sweeps; FP64 (96 frames) 8 chunks of 48,
        // three full rounds, 5.38 ms, against 6.19 for 2 chunks of 192
        // (one round with 64 CUs idle); the model orders all seven chunk
        // sizes tried at both precisions as measured.
        long long best_cost = -1;
        for (int nch = 1; nch <= c.nz; ++nch) {
            const int ch = (c.nz + nch - 1) / nch;
            if (ch < 2 * c.r) break;
            const long long nt = (long long)tiles_y * ((c.nz + ch - 1) / ch);
            const long long cost = (nt + cus - 1) / cus * (ch + 2 * c.r);
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                chunk = ch;
            }
        }
        if (chunk <= 0) chunk = c.nz;
    }
    c.kind = 4;
    c.cx = cx;
    c.cy = 1;
    c.hs = hs;
    c.wx = wx;
    c.wy = wy;
    c.bx = c.nx;
    c.by = wy;
    c.bz = chunk;
    c.zchunk = chunk;
    c.gx = tiles_y * ((c.nz + chunk - 1) / chunk);
    c.gy = c.gz = 1;
    return ST_OK;
}

st_status resolve_rows(Cfg &c) { return rows_resolve_mode(c, RM_SWEEP, false); }

// sweeps > 1: one cooperative launch runs them all, the buffers swapped
// after each (the grid is capped at the co-resident workgroups; the tile loop
// strides over the rest)
static st_status rows_launch(const Cfg &c, const void *fn, const void *in, void *out, const uint16_t *aux,
                             hipStream_t s, int sweeps = 1, uint16_t *eout = nullptr, float *xout = nullptr)
{
    RowsArgs a;
    a.eout = eout;
    a.xout = xout;
    a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
    a.sz = (long long)c.nx * c.ny;
    a.wx = c.wx;
    a.wy = c.wy;
    a.tiles_y = (c.ny + c.wy - 1) / c.wy;
    a.zchunk = c.zchunk;
    a.ntiles = a.tiles_y * ((c.nz + c.zchunk - 1) / c.zchunk);
    a.nt = 1;
    a.sweeps = sweeps;
    a.eq = c.shell_eq;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
    if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(rows LDS)");
    const int nthreads = 64 * c.wx * c.wy;
    int grid = a.ntiles;
    if (sweeps > 1) {
        int per_cu = 0, dev = 0, cus = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nthreads, c.lds);
        if (e != hipSuccess) return hip_check(e, "occupancy query (rows)");
        if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_check(e, "hipGetDevice");
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_check(e, "CU count");
        if (per_cu < 1 || cus < 1) return fail(ST_ELDS, "row kernel does not fit on a CU (lds %d)", c.lds);
        if ((long long)per_cu * cus < grid) grid = per_cu * cus;
    }
    double kd = c.op == OP_THESIS ? (double)(2 * c.dims * c.r) : c.weight;
    float kf = (float)kd;
    void *karg = c.dtype == ST_F64 ? static_cast<void *>(&kd) : static_cast<void *>(&kf);
    void *args[] = {&in, &out, &aux, &a, karg};
    if (sweeps > 1) {
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(nthreads), args, c.lds, s);
        if (e == hipErrorIllegalState) {
            // the device takes no cooperative launch from this process now:
            // the same sweeps, one launch each, the buffers swapped between
            (void)hipGetLastError();
            const void *a_in = in;
            void *a_out = out;
            for (int i = 0; i < sweeps; ++i) {
                st_status st = rows_launch(c, fn, a_in, a_out, aux, s, 1);
                if (st != ST_OK) return st;
                void *t = const_cast<void *>(a_in);
                a_in = a_out;
                a_out = t;
            }
            return ST_OK;
        }
        if (e != hipSuccess) return hip_check(e, "hipLaunchCooperativeKernel(rows)");
    } else {
        e = hipLaunchKernel(fn, dim3(grid), dim3(nthreads), args, c.lds, s);
        if (e != hipSuccess) return hip_check(e, "rows launch");
    }
    return ST_OK;
}

st_status launch_rows(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    int threads = 0;
    const void *fn = rows_kernel(c, RM_SWEEP, false, &threads);
    if (!fn || 64 * c.wx * c.wy > threads) return fail(ST_EUNSUPPORTED, "row-streaming kernel missing");
    return rows_launch(c, fn, in, out, nullptr, s);
}

// COOP variant, row form: the row kernel's tiles, every sweep in one
// cooperative launch with a grid barrier between sweeps (or one launch per
// sweep with launch_per_pass).
st_status resolve_rows_coop(Cfg &c)
{
    if (!rows_eligible(c))
        return fail(ST_EUNSUPPORTED, "COOP row form needs a 3D F32/F64 grid, radius 1..4 and nx a multiple "
                                     "of 4 whose rows fit a workgroup with its halo strips");
    st_status s = rows_resolve_mode(c, RM_SWEEP, false);
    if (s != ST_OK) return s;
    c.kind = 5;
    c.tsteps = 1;
    return ST_OK;
}

st_status run_rows_coop(const Cfg &c0, void *u1, void *u2, hipStream_t s, int *in_u2)
{
    *in_u2 = c0.iters & 1;
    if (c0.iters <= 0) return ST_OK;
    Cfg c = c0;
    int threads = 0;
    const void *fn = rows_kernel(c, RM_SWEEP, false, &threads);
    if (!fn || 64 * c.wx * c.wy > threads) return fail(ST_EUNSUPPORTED, "row-streaming kernel missing");
    st_status st = ST_OK;
    if (c.iters > 1) st = rows_shell_check(c, u1, u2, s, &c.shell_eq);
    if (st != ST_OK) return st;
    if (!c.per_pass) return rows_launch(c, fn, u1, u2, nullptr, s, c.iters);
    void *a = u1, *b = u2;
    for (int i = 0; i < c.iters && st == ST_OK; ++i) {
        st = rows_launch(c, fn, a, b, nullptr, s);
        void *t = a; a = b; b = t;
    }
    return st;
}

// Shell comparison of one grid: k_flag_one sets the flag, any block that
// finds a shell cell whose bits differ clears it.
__global__ __launch_bounds__(64) void k_flag_one(int *__restrict__ flag)
{
    if (threadIdx.x == 0) *flag = 1;
}

template <typename W>
__global__ __launch_bounds__(256) void k_grid_shell_eq(const W *__restrict__ u1, const W *__restrict__ u2, int nx,
                                                       int ny, int nz, int R, int *__restrict__ flag)
{
    const unsigned n1 = (unsigned)shell_full_cells(nx, ny, nz, R), n = (unsigned)shell_cells(nx, ny, nz, R);
    int ok = 1;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n; i += 256u * gridDim.x) {
        const long long idx = shell_cell(i, n1, nx, ny, nz, R);
        ok &= u1[idx] == u2[idx];
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0 && !ok) *flag = 0;
}

st_status rows_shell_check(const Cfg &c, const void *u1, const void *u2, hipStream_t s, const int **flag)
{
    const long long n = shell_cells(c.nx, c.ny, c.nz, c.r);
    if (n >= (1LL << 31)) {  // not compared: interior-only stores
        *flag = nullptr;
        return ST_OK;
    }
    void *ws = nullptr;
    st_status st = workspace_slot(2, sizeof(int), s, &ws);
    if (st != ST_OK) return st;
    int *f = static_cast<int *>(ws);
    const int blocks = (int)std::max<long long>(1, std::min<long long>(2048, (n + 2047) / 2048));
    hipLaunchKernelGGL(k_flag_one, dim3(1), dim3(64), 0, s, f);
    if (c.dtype == ST_F64)
        hipLaunchKernelGGL(k_grid_shell_eq<unsigned long long>, dim3(blocks), dim3(256), 0, s,
                           static_cast<const unsigned long long *>(u1), static_cast<const unsigned long long *>(u2),
                           c.nx, c.ny, c.nz, c.r, f);
    else
        hipLaunchKernelGGL(k_grid_shell_eq<unsigned>, dim3(blocks), dim3(256), 0, s, static_cast<const unsigned *>(u1),
                           static_cast<const unsigned *>(u2), c.nx, c.ny, c.nz, c.r, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_check(e, "shell comparison launch");
    *flag = f;
    return ST_OK;
}

// The 16-bit modes resolve their own layout (their register need differs
// from the sweep's by the defect row).
bool rows_mixed_fused_ok(const Cfg &c)
{
    if (c.dims != 3) return false;
    for (int mode : {RM_CORR, RM_DEFECT, RM_FOLD}) {
        Cfg f = c;
        f.dtype = ST_F32;
        f.op = OP_JACOBI;
        f.heuristic = 1;
        if (rows_resolve_mode(f, mode, c.dtype == ST_F16) != ST_OK || !f.rowc) return false;
    }
    return true;
}

st_status launch_rows_defect(const Cfg &c, bool f16, const float *x, const uint16_t *e_old, float *x_out,
                             uint16_t *d, uint16_t *e_first, hipStream_t s)
{
    Cfg rc = c;
    rc.dtype = ST_F32;
    rc.op = OP_JACOBI;
    rc.heuristic = 1;
    const int mode = e_old ? RM_FOLD : RM_DEFECT;
    st_status st = rows_resolve_mode(rc, mode, f16);
    if (st != ST_OK) return st;
    if (!rc.rowc) return fail(ST_EUNSUPPORTED, "fused defect sweep needs the consecutive row layout");
    int threads = 0;
    const void *fn = rows_kernel(rc, mode, f16, &threads);
    if (!fn || 64 * rc.wx * rc.wy > threads) return fail(ST_EUNSUPPORTED, "row-streaming defect kernel missing");
    return rows_launch(rc, fn, x, d, e_old, s, 1, e_first, x_out);
}

bool rows_mixed_eligible(const Cfg &c)
{
    Cfg f = c;
    f.dtype = ST_F32;
    f.op = OP_JACOBI;
    f.heuristic = 1;
    Cfg g = f;
    return f.dims == 3 && rows_resolve_mode(f, RM_CORR, c.dtype == ST_F16) == ST_OK &&
           rows_resolve_mode(g, RM_DEFECT, c.dtype == ST_F16) == ST_OK;
}

// `out` must have a zero shell (edge quads write zeros there).
st_status launch_rows_mixed(const Cfg &c, int defect, bool f16, const void *in, void *out, const uint16_t *d,
                            hipStream_t s)
{
    Cfg rc = c;
    rc.dtype = ST_F32;
    rc.op = OP_JACOBI;
    rc.heuristic = 1;
    const int mode = defect ? RM_DEFECT : RM_CORR;
    st_status st = rows_resolve_mode(rc, mode, f16);
    if (st != ST_OK) return st;
    int threads = 0;
    const void *fn = rows_kernel(rc, mode, f16, &threads);
    if (!fn || 64 * rc.wx * rc.wy > threads) return fail(ST_EUNSUPPORTED, "row-streaming mixed kernel missing");
    return rows_launch(rc, fn, in, out, d, s);
}

}  // namespace st
