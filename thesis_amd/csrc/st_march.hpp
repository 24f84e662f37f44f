// Register-queue march of one wave tile, shared by the single-grid
// SMEM_REGISTER kernels (sweep_march.hip) and the batched kernels
// (batched.hip).  See sweep_march.hip for the design.
#pragma once

#include "st_device.hpp"

namespace st {

template <typename T, int N> struct VecT;
template <> struct VecT<float, 1> { typedef float type; };
template <> struct VecT<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct VecT<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecT<double, 1> { typedef double type; };
template <> struct VecT<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T, int N>
__device__ __forceinline__ void vload(const T *p, T (&d)[N])
{
    if constexpr (N == 1) {
        d[0] = *p;
    } else {
        typedef typename VecT<T, N>::type V;
        const V v = *reinterpret_cast<const V *>(p);
#pragma unroll
        for (int c = 0; c < N; ++c) d[c] = v[c];
    }
}

template <typename T, int N>
__device__ __forceinline__ void vstore(T *p, const T (&s)[N], bool nt)
{
    if constexpr (N == 1) {
        if (nt) __builtin_nontemporal_store(s[0], p);
        else *p = s[0];
    } else {
        typedef typename VecT<T, N>::type V;
        V v;
#pragma unroll
        for (int c = 0; c < N; ++c) v[c] = s[c];
        if (nt) __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
        else *reinterpret_cast<V *>(p) = v;
    }
}


// One wave streams one tile: x range [tx*64*CX, +64*CX), rows
// [ty*TY, +TY) (3D), march-axis chunk [tz*zchunk, +zchunk).
// D = 3: rows are y, march axis z.   D = 2: one row (TY = 1), march axis y.
template <int OP, int D, int R, int CX, int TY, typename T>
__device__ __forceinline__ void march_tile(const T *__restrict__ in, T *__restrict__ out,
                                           const Geom &g, int tx, int ty, int tz, int zchunk,
                                           T kc, int nt)
{
    static_assert(D == 3 || TY == 1, "2D march uses one row per lane");
    static_assert(2 * R * TY <= 64, "edge cells must fit one load per lane");
    constexpr int Q = 2 * R + 2;  // register ring depth (planes)
    const int lane = threadIdx.x & 63;
    const int xw0 = tx * 64 * CX;  // first x of the wave
    const int x0 = xw0 + lane * CX;
    const int j0 = (D == 3) ? ty * TY : 0;
    const long long ms = (D == 3) ? g.sz : g.sy;  // march-axis stride
    const int mn = (D == 3) ? g.nz : g.ny;        // march-axis extent
    const int kb = max(tz * zchunk, R);
    const int ke = min(tz * zchunk + zchunk, mn - R);
    if (kb >= ke) return;

    // Clamped (always in-bounds) load coordinates; clamped values are only
    // ever consumed by cells that are not written.
    const int xl = min(x0, g.nx - CX);
    int rowoff[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) rowoff[r] = (D == 3) ? min(j0 + r, g.ny - 1) * (int)g.sy : 0;
    int hlo_off[R], hhi_off[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        hlo_off[m] = (D == 3) ? max(j0 - R + m, 0) * (int)g.sy : 0;
        hhi_off[m] = (D == 3) ? min(j0 + TY + m, g.ny - 1) * (int)g.sy : 0;
    }
    // Edge cell owned by this lane: lanes [0, R*TY) hold the R cells left of
    // the wave for each row, lanes [R*TY, 2*R*TY) the R cells right of it.
    const bool has_edge = lane < 2 * R * TY;
    long long edge_off;
    {
        const int t = lane < R * TY ? lane : lane - R * TY;
        const int r = t / R, m = t % R;
        int ex = lane < R * TY ? xw0 - R + m : xw0 + 64 * CX + m;
        ex = min(max(ex, 0), g.nx - 1);
        const int er = (D == 3) ? min(j0 + min(r, TY - 1), g.ny - 1) : 0;
        edge_off = has_edge ? (long long)er * g.sy + ex : 0;
    }

    T q[Q][TY][CX];
    T hlo[R][CX], hhi[R][CX], nhlo[R][CX], nhhi[R][CX];
    T edge, nedge;

    auto load_plane = [&](int kp, T (&dst)[TY][CX]) {
        const T *p = in + (long long)kp * ms + xl;
#pragma unroll
        for (int r = 0; r < TY; ++r) vload<T, CX>(p + rowoff[r], dst[r]);
    };
    auto load_halo = [&](int kp, T (&lo)[R][CX], T (&hi)[R][CX], T &e) {
        const T *p = in + (long long)kp * ms;
        if constexpr (D == 3) {
#pragma unroll
            for (int m = 0; m < R; ++m) {
                vload<T, CX>(p + hlo_off[m] + xl, lo[m]);
                vload<T, CX>(p + hhi_off[m] + xl, hi[m]);
            }
        }
        e = has_edge ? p[edge_off] : (T)0;
    };

    // prologue: planes kb-R .. kb+R into ring slots 0 .. 2R
#pragma unroll
    for (int p = 0; p < 2 * R + 1; ++p) load_plane(kb - R + p, q[p]);
    load_halo(kb, hlo, hhi, edge);

    for (int k = kb; k < ke; k += Q) {
#pragma unroll
        for (int ph = 0; ph < Q; ++ph) {
            const int kk = k + ph;
            if (kk >= ke) break;
            // prefetch: plane kk+R+1 into the slot of plane kk-R-1, halo of kk+1
            load_plane(min(kk + R + 1, mn - 1), q[(ph + 2 * R + 1) % Q]);
            load_halo(min(kk + 1, mn - 1), nhlo, nhhi, nedge);

            T(&cur)[TY][CX] = q[(ph + R) % Q];
            T *orow = out + (long long)kk * ms;
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                // x halo of this row: L[m] at x0-R+m, Rv[m] at x0+CX+m
                T L[R], Rv[R];
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const int rel = m - R;  // negative
                    const int off = floor_div(rel, CX);
                    const int comp = rel - off * CX;
                    const int src = lane + off;
                    const int p = lane * CX + rel;
                    T own = cur[r][0];
#pragma unroll
                    for (int c = 0; c < CX; ++c)
                        if (c == comp) own = cur[r][c];
                    const T va = lane_read(own, src);
                    const T vb = lane_read(edge, r * R + p + R);
                    L[m] = (src >= 0) ? va : vb;
                }
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const int rel = CX + m;
                    const int off = rel / CX;
                    const int comp = rel % CX;
                    const int src = lane + off;
                    const int p = lane * CX + rel - 64 * CX;
                    T own = cur[r][0];
#pragma unroll
                    for (int c = 0; c < CX; ++c)
                        if (c == comp) own = cur[r][c];
                    const T va = lane_read(own, src);
                    const T vb = lane_read(edge, R * TY + r * R + p);
                    Rv[m] = (src < 64) ? va : vb;
                }
                T res[CX];
#pragma unroll
                for (int c = 0; c < CX; ++c) {
                    T u = 0;
#pragma unroll
                    for (int d = R; d >= 1; --d) u += (c - d >= 0) ? cur[r][c - d] : L[R + c - d];
#pragma unroll
                    for (int d = 1; d <= R; ++d) u += (c + d < CX) ? cur[r][c + d] : Rv[c + d - CX];
                    if constexpr (D == 3) {
#pragma unroll
                        for (int d = R; d >= 1; --d) u += (r - d >= 0) ? cur[r - d][c] : hlo[R + r - d][c];
#pragma unroll
                        for (int d = 1; d <= R; ++d) u += (r + d < TY) ? cur[r + d][c] : hhi[r + d - TY][c];
                    }
#pragma unroll
                    for (int d = R; d >= 1; --d) u += q[(ph + R - d) % Q][r][c];
#pragma unroll
                    for (int d = 1; d <= R; ++d) u += q[(ph + R + d) % Q][r][c];
                    res[c] = finish<OP>(u, cur[r][c], kc);
                }
                const int j = (D == 3) ? j0 + r : kk;
                if (!interior1(j, g.ny, R)) continue;
                T *po = orow + ((D == 3) ? (long long)j * g.sy : 0) + x0;
                if (x0 >= R && x0 + CX <= g.nx - R) {
                    vstore<T, CX>(po, res, nt);
                } else {
#pragma unroll
                    for (int c = 0; c < CX; ++c)
                        if (interior1(x0 + c, g.nx, R)) po[c] = res[c];
                }
            }
#pragma unroll
            for (int m = 0; m < R; ++m)
#pragma unroll
                for (int c = 0; c < CX; ++c) {
                    hlo[m][c] = nhlo[m][c];
                    hhi[m][c] = nhhi[m][c];
                }
            edge = nedge;
        }
    }
}

}  // namespace st
