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
#include "st_march.hpp"

namespace st {

struct MarchArgs {
    Geom g;
    int zchunk;       // planes (3D) or rows (2D) per wave tile
    int tiles_x;      // wave tiles along x
    int tiles_y;      // wave tiles along y (3D), 1 in 2D
    int tiles_total;  // tiles_x * tiles_y * chunks
    int nblocks;      // workgroups launched
    int nt;           // non-temporal stores
};

template <int OP, int D, int R, int CX, int TY, typename T>
__global__ __launch_bounds__(256) void k_march(const T *__restrict__ in, T *__restrict__ out,
                                               MarchArgs a, T kc)
{
    const int wpb = blockDim.x >> 6;
    const unsigned b = xcd_remap(blockIdx.x, (unsigned)a.nblocks);
    const int wt = (int)(b * wpb + (threadIdx.x >> 6));
    if (wt >= a.tiles_total) return;
    const int tx = wt % a.tiles_x;
    const int rest = wt / a.tiles_x;
    march_tile<OP, D, R, CX, TY, T>(in, out, a.g, tx, rest % a.tiles_y, rest / a.tiles_y, a.zchunk,
                                    kc, a.nt);
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
    ME(3, float, 9, 1, 1), ME(3, float, 10, 1, 1), ME(3, float, 11, 1, 1), ME(3, float, 12, 1, 1),
    ME(3, float, 13, 1, 1), ME(3, float, 14, 1, 1), ME(3, float, 15, 1, 1), ME(3, float, 16, 1, 1),
    // 3D double
    ME(3, double, 1, 1, 1), ME(3, double, 1, 1, 2), ME(3, double, 1, 1, 4),
    ME(3, double, 1, 2, 1), ME(3, double, 1, 2, 2), ME(3, double, 1, 2, 4),
    ME(3, double, 2, 1, 1), ME(3, double, 2, 2, 1), ME(3, double, 2, 2, 2),
    ME(3, double, 4, 1, 1), ME(3, double, 4, 2, 1), ME(3, double, 4, 2, 2),
    ME(3, double, 3, 1, 1), ME(3, double, 5, 1, 1), ME(3, double, 6, 1, 1), ME(3, double, 7, 1, 1),
    ME(3, double, 8, 1, 1), ME(3, double, 9, 1, 1), ME(3, double, 10, 1, 1), ME(3, double, 11, 1, 1),
    ME(3, double, 12, 1, 1), ME(3, double, 13, 1, 1), ME(3, double, 14, 1, 1), ME(3, double, 15, 1, 1),
    ME(3, double, 16, 1, 1),
    // 2D
    ME(2, float, 1, 1, 1), ME(2, float, 1, 2, 1), ME(2, float, 1, 4, 1),
    ME(2, float, 2, 1, 1), ME(2, float, 2, 4, 1), ME(2, float, 3, 1, 1), ME(2, float, 3, 4, 1),
    ME(2, float, 4, 1, 1), ME(2, float, 4, 4, 1), ME(2, float, 5, 1, 1), ME(2, float, 6, 1, 1),
    ME(2, float, 7, 1, 1), ME(2, float, 8, 1, 1), ME(2, float, 9, 1, 1), ME(2, float, 10, 1, 1),
    ME(2, float, 11, 1, 1), ME(2, float, 12, 1, 1), ME(2, float, 13, 1, 1), ME(2, float, 14, 1, 1),
    ME(2, float, 15, 1, 1), ME(2, float, 16, 1, 1),
    ME(2, double, 1, 1, 1), ME(2, double, 1, 2, 1), ME(2, double, 2, 1, 1),
    ME(2, double, 2, 2, 1), ME(2, double, 4, 1, 1), ME(2, double, 4, 2, 1),
    ME(2, double, 3, 1, 1), ME(2, double, 5, 1, 1), ME(2, double, 6, 1, 1), ME(2, double, 7, 1, 1),
    ME(2, double, 8, 1, 1), ME(2, double, 9, 1, 1), ME(2, double, 10, 1, 1), ME(2, double, 11, 1, 1),
    ME(2, double, 12, 1, 1), ME(2, double, 13, 1, 1), ME(2, double, 14, 1, 1), ME(2, double, 15, 1, 1),
    ME(2, double, 16, 1, 1),
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
    // form: 1 register march, 2 row streaming; the library's choice is the
    // row-streaming form wherever it applies (3D, radius 1..4: 599 vs 582
    // GCells/s at 512^3 R=1, 2.9x at 384^3 R=4), the march elsewhere
    {
        const bool heur0 = c.heuristic || c.bx <= 0;
        if (c.form == 2 || (c.form == 0 && heur0 && rows_eligible(c))) {
            if (heur0) c.heuristic = 1;
            return resolve_rows(c);
        }
        if (c.form > 2) return fail(ST_EINVAL, "SMEM_REGISTER forms are 0 (auto), 1 (march), 2 (rows)");
    }
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
        return fail(ST_EUNSUPPORTED, "SMEM_REGISTER is built for radius 1..16; radius %d requested", c.r);
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
    if (c.kind == 4) return launch_rows(c, in, out, s);
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
