// SMEM_REGISTER family: register-queue march along the slowest axis.
//
// Replaces smem_register_{2d,3d} and their unrolled forms
// (src/stencil_kernel_smem_register.cu:8-116, src/prefetch_smem.cu:151-258,
// src/stencils.cu:71-83).  The reference keeps one (2R+1)-deep column per
// thread for a single output point; here every wave streams along the march
// axis (z in 3D, y in 2D) and keeps a ring of 2R+2 planes in registers, so
// each input value is read from HBM once per sweep:
//
//   * a wave owns a tile of 64*CX cells along x (CX consecutive cells per
//     lane, one 64-lane vector load per row and plane: 256 B .. 1 KiB fully
//     coalesced) and TY rows along y (3D only);
//   * z neighbours come from the register ring, y neighbours from the wave's
//     own rows plus R halo rows above and below, x neighbours from the
//     neighbouring lanes through ds_bpermute (no LDS allocation, no barrier),
//     and the R cells beyond each end of the wave's x range from one packed
//     edge load per plane;
//   * waves never synchronise, so latency is hidden by the next plane's
//     loads being in flight while the current plane is computed;
//   * the workgroup id is remapped so that each XCD processes a contiguous
//     range of tiles (its y neighbours hit in the same L2).
// Summation order is the reference's, so results are bit-identical to the
// other families and to the CPU oracle.
#include "st_device.hpp"
#include "st_internal.hpp"

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

struct MarchArgs {
    Geom g;
    int zchunk;       // planes (3D) or rows (2D) per wave tile
    int tiles_x;      // wave tiles along x
    int tiles_y;      // wave tiles along y (3D), 1 in 2D
    int tiles_total;  // tiles_x * tiles_y * chunks
    int nblocks;      // workgroups launched
    int nt;           // non-temporal stores
};

// D = 3: rows are y, march axis z.   D = 2: one row (TY = 1), march axis y.
template <int OP, int D, int R, int CX, int TY, typename T>
__global__ __launch_bounds__(256) void k_march(const T *__restrict__ in, T *__restrict__ out,
                                               MarchArgs a, T kc)
{
    static_assert(D == 3 || TY == 1, "2D march uses one row per lane");
    static_assert(2 * R * TY <= 64, "edge cells must fit one load per lane");
    constexpr int Q = 2 * R + 2;  // register ring depth (planes)
    const Geom &g = a.g;
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const unsigned b = xcd_remap(blockIdx.x, (unsigned)a.nblocks);
    const int wt = (int)(b * wpb + (threadIdx.x >> 6));
    if (wt >= a.tiles_total) return;
    const int tx = wt % a.tiles_x;
    const int rest = wt / a.tiles_x;
    const int ty = rest % a.tiles_y;
    const int tz = rest / a.tiles_y;

    const int xw0 = tx * 64 * CX;  // first x of the wave
    const int x0 = xw0 + lane * CX;
    const int j0 = (D == 3) ? ty * TY : 0;
    const long long ms = (D == 3) ? g.sz : g.sy;  // march-axis stride
    const int mn = (D == 3) ? g.nz : g.ny;        // march-axis extent
    const int kb = max(tz * a.zchunk, R);
    const int ke = min(tz * a.zchunk + a.zchunk, mn - R);
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
                    vstore<T, CX>(po, res, a.nt);
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

// ---------------------------------------------------------------- instantiation table

typedef void (*MarchFn)(const void *, void *, const MarchArgs &, double, dim3, dim3,
                        hipStream_t);

template <int OP, int D, int R, int CX, int TY, typename T>
static void march_launch(const void *in, void *out, const MarchArgs &a, double kc, dim3 grid,
                         dim3 block, hipStream_t s)
{
    hipLaunchKernelGGL((k_march<OP, D, R, CX, TY, T>), grid, block, 0, s,
                       static_cast<const T *>(in), static_cast<T *>(out), a, (T)kc);
}

struct MarchEntry {
    int d, f64, r, cx, ty;
    MarchFn fn[2];  // indexed by operator
};

#define ME(D, T, R, CX, TY)                                                        \
    {D, sizeof(T) == 8, R, CX, TY,                                                 \
     {&march_launch<OP_THESIS, D, R, CX, TY, T>, &march_launch<OP_JACOBI, D, R, CX, TY, T>}}

static const MarchEntry kMarch[] = {
    // 3D float
    ME(3, float, 1, 1, 1), ME(3, float, 1, 1, 2), ME(3, float, 1, 1, 4), ME(3, float, 1, 1, 8),
    ME(3, float, 1, 2, 1), ME(3, float, 1, 2, 2), ME(3, float, 1, 2, 4), ME(3, float, 1, 2, 8),
    ME(3, float, 1, 4, 1), ME(3, float, 1, 4, 2), ME(3, float, 1, 4, 4), ME(3, float, 1, 4, 8),
    ME(3, float, 2, 1, 1), ME(3, float, 2, 1, 2), ME(3, float, 2, 1, 4),
    ME(3, float, 2, 4, 1), ME(3, float, 2, 4, 2), ME(3, float, 2, 4, 4),
    ME(3, float, 3, 1, 1), ME(3, float, 3, 1, 2), ME(3, float, 3, 4, 1), ME(3, float, 3, 4, 2),
    ME(3, float, 4, 1, 1), ME(3, float, 4, 1, 2), ME(3, float, 4, 4, 1), ME(3, float, 4, 4, 2),
    ME(3, float, 5, 1, 1), ME(3, float, 5, 1, 2), ME(3, float, 6, 1, 1), ME(3, float, 6, 1, 2),
    ME(3, float, 7, 1, 1), ME(3, float, 7, 1, 2), ME(3, float, 8, 1, 1), ME(3, float, 8, 1, 2),
    // 3D double
    ME(3, double, 1, 1, 1), ME(3, double, 1, 1, 2), ME(3, double, 1, 1, 4),
    ME(3, double, 1, 2, 1), ME(3, double, 1, 2, 2), ME(3, double, 1, 2, 4),
    ME(3, double, 2, 1, 1), ME(3, double, 2, 2, 1), ME(3, double, 2, 2, 2),
    ME(3, double, 4, 1, 1), ME(3, double, 4, 2, 1), ME(3, double, 4, 2, 2),
    // 2D
    ME(2, float, 1, 1, 1), ME(2, float, 1, 2, 1), ME(2, float, 1, 4, 1),
    ME(2, float, 2, 1, 1), ME(2, float, 2, 4, 1), ME(2, float, 3, 1, 1), ME(2, float, 3, 4, 1),
    ME(2, float, 4, 1, 1), ME(2, float, 4, 4, 1), ME(2, float, 5, 1, 1), ME(2, float, 6, 1, 1),
    ME(2, float, 7, 1, 1), ME(2, float, 8, 1, 1),
    ME(2, double, 1, 1, 1), ME(2, double, 1, 2, 1), ME(2, double, 2, 1, 1),
    ME(2, double, 2, 2, 1), ME(2, double, 4, 1, 1), ME(2, double, 4, 2, 1),
};
#undef ME

static const MarchEntry *find_march(int d, int f64, int r, int cx, int ty)
{
    for (const MarchEntry &e : kMarch)
        if (e.d == d && e.f64 == f64 && e.r == r && e.cx == cx && e.ty == ty) return &e;
    return nullptr;
}

// Largest instantiated (cx, ty) not above the request.
static const MarchEntry *best_march(int d, int f64, int r, int cx, int ty)
{
    for (int c = cx; c >= 1; c /= 2)
        for (int t = ty; t >= 1; t /= 2)
            if (const MarchEntry *e = find_march(d, f64, r, c, t)) return e;
    return nullptr;
}

st_status resolve_march(Cfg &c)
{
    const int f64 = c.dtype == ST_F64;
    int cx = c.cx > 0 ? c.cx : 1;
    int ty = c.dims == 3 ? (c.cy > 0 ? c.cy : 0) : 1;
    const bool heur = c.heuristic || c.bx <= 0;
    if (heur) {
        // Full-width 1 KiB row segments for FP32 R=1 (float4 per lane);
        // narrower tiles for large radii keep the register ring small.
        if (f64) cx = 2;
        else cx = c.r <= 2 ? 4 : (c.r <= 4 ? 4 : 1);
        if (ty == 0) ty = c.dims == 3 ? (c.r == 1 ? 4 : (c.r <= 2 ? 2 : 1)) : 1;
    }
    if (ty == 0) ty = c.dims == 3 ? 2 : 1;
    while (cx > 1 && c.nx % cx) cx /= 2;
    if (cx > 4) cx = 4;
    const MarchEntry *e = best_march(c.dims, f64, c.r, cx, ty);
    if (!e)
        return fail(ST_EUNSUPPORTED, "SMEM_REGISTER is built for radius 1..8 (FP32) and 1,2,4 "
                                     "(FP64); radius %d requested", c.r);
    c.cx = e->cx;
    c.cy = e->ty;
    // Workgroup: block_x threads (a multiple of 64) along x, block_y waves of rows.
    int bx = heur ? 256 : c.bx;
    if (bx % 64) return fail(ST_EINVAL, "SMEM_REGISTER needs block_x to be a multiple of 64, got %d", bx);
    int by = heur ? 1 : (c.by > 0 ? c.by : 1);
    if (bx * by > kMaxThreads) return fail(ST_EINVAL, "block %dx%d exceeds 1024 threads", bx, by);
    const int wpb = bx / 64 * by;
    c.bx = 64 * wpb;
    c.by = 1;
    const int tiles_x = (c.nx + 64 * c.cx - 1) / (64 * c.cx);
    const int tiles_y = c.dims == 3 ? (c.ny + c.cy - 1) / c.cy : 1;
    const int mn = c.dims == 3 ? c.nz : c.ny;
    int chunk = c.dims == 3 && !heur && c.bz > 1 ? c.bz : 0;
    if (chunk <= 0) {
        // Enough wave tiles to fill 256 CUs x 8 waves, chunks of >= 16 planes.
        const long long per_chunk = (long long)tiles_x * tiles_y;
        long long chunks = (2048 + per_chunk - 1) / per_chunk;
        if (chunks < 1) chunks = 1;
        chunk = (int)((mn + chunks - 1) / chunks);
        if (chunk < 16) chunk = 16;
        if (chunk > mn) chunk = mn;
    }
    c.zchunk = chunk;
    const long long chunks = (mn + chunk - 1) / chunk;
    const long long tiles = (long long)tiles_x * tiles_y * chunks;
    if (tiles > (1LL << 30)) return fail(ST_EINVAL, "too many tiles");
    c.gx = (int)((tiles + wpb - 1) / wpb);
    c.gy = 1;
    c.gz = 1;
    c.bz = 1;
    c.lds = 0;
    c.tsteps = 1;
    return ST_OK;
}

st_status launch_march(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    const MarchEntry *e = find_march(c.dims, c.dtype == ST_F64, c.r, c.cx, c.cy);
    if (!e) return fail(ST_EUNSUPPORTED, "no SMEM_REGISTER kernel for r=%d cx=%d ty=%d", c.r, c.cx, c.cy);
    const int es = c.dtype == ST_F64 ? 8 : 4;
    if (c.cx > 1 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) %
                     (uintptr_t)(es * c.cx)))
        return fail(ST_EINVAL, "SMEM_REGISTER with coarsen_x %d needs %d-byte aligned buffers",
                    c.cx, es * c.cx);
    MarchArgs a;
    a.g.nx = c.nx; a.g.ny = c.ny; a.g.nz = c.nz; a.g.r = c.r;
    a.g.sy = c.nx; a.g.sz = (long long)c.nx * c.ny;
    a.zchunk = c.zchunk;
    a.tiles_x = (c.nx + 64 * c.cx - 1) / (64 * c.cx);
    a.tiles_y = c.dims == 3 ? (c.ny + c.cy - 1) / c.cy : 1;
    const int mn = c.dims == 3 ? c.nz : c.ny;
    a.tiles_total = a.tiles_x * a.tiles_y * ((mn + c.zchunk - 1) / c.zchunk);
    a.nblocks = c.gx;
    a.nt = 1;
    const double kc = c.op == OP_THESIS ? (double)(2 * c.dims * c.r) : c.weight;
    e->fn[c.op](in, out, a, kc, dim3(c.gx), dim3(c.bx), s);
    return hip_check(hipGetLastError(), "march launch");
}

}  // namespace st
