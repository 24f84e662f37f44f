// Single-sweep kernels that keep the reference's thread-per-cell mapping:
//   BASE         all operands from HBM / cache
//                (replaces base_{2d,3d}, base_unroll_*: src/stencil_kernel_base.cu:5-65)
//   SMEM         LDS tile without halo, halo operands from global
//                (replaces smem_*, smem_unroll_*: src/stencil_kernel_smem.cu:8-106,
//                 src/stencils.cu:21-55, src/stencil_accumulators.cu:18-32,112-140)
//   SMEM_PADDED  LDS tile with an R-deep halo, loaded cooperatively by the
//                whole workgroup with row-contiguous (coalesced) reads
//                (replaces smem_padded_*: src/stencil_kernel_smem_padded.cu:8-119,
//                 src/prefetch_smem.cu:7-149)
// Coarsening follows the reference: cell s of a thread is s*blockDim.x to
// the right of its first cell, so every coarsening step is one coalesced
// 64-lane row segment.  Kernels are templated on the operator, the
// dimensionality and the radius (RS = 0 means the radius is read from
// Geom::r at run time).
#include "st_device.hpp"
#include "st_internal.hpp"

namespace st {

// ---------------------------------------------------------------- BASE

template <int OP, int D, int RS, typename T>
__global__ void k_base(const T *__restrict__ in, T *__restrict__ out, Geom g, int cx, T kc)
{
    const int R = RS ? RS : g.r;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int k = (D == 3) ? (int)(threadIdx.z + blockIdx.z * blockDim.z) : 0;
    if (!interior1(j, g.ny, R)) return;
    if (D == 3 && !interior1(k, g.nz, R)) return;
    const int ib = threadIdx.x + blockIdx.x * blockDim.x * cx;
    for (int s = 0; s < cx; ++s) {
        const int i = ib + s * (int)blockDim.x;
        if (!interior1(i, g.nx, R)) continue;
        const long long idx = i + j * g.sy + k * g.sz;
        T u = 0;
#pragma unroll
        for (int d = R; d >= 1; --d) u += in[idx - d];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += in[idx + d];
#pragma unroll
        for (int d = R; d >= 1; --d) u += in[idx - d * g.sy];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += in[idx + d * g.sy];
        if constexpr (D == 3) {
#pragma unroll
            for (int d = R; d >= 1; --d) u += in[idx - d * g.sz];
#pragma unroll
            for (int d = 1; d <= R; ++d) u += in[idx + d * g.sz];
        }
        out[idx] = finish<OP>(u, in[idx], kc);
    }
}

// ---------------------------------------------------------------- SMEM (hybrid)

template <int OP, int D, int RS, typename T>
__global__ void k_smem(const T *__restrict__ in, T *__restrict__ out, Geom g, int cx, T kc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // explicit LDS and global views: the per-neighbour "tile or global"
    // choice then stays two loads and a value select instead of one flat
    // load through a selected pointer
    __attribute__((address_space(3))) T *tile = (__attribute__((address_space(3))) T *)lds_raw;
    const __attribute__((address_space(1))) T *gin = (const __attribute__((address_space(1))) T *)in;
    const int R = RS ? RS : g.r;
    const int bx = blockDim.x, by = blockDim.y, bz = (D == 3) ? blockDim.z : 1;
    const int tx = threadIdx.x, ty = threadIdx.y, tz = (D == 3) ? threadIdx.z : 0;
    const int TX = bx * cx;
    const int ib = tx + blockIdx.x * TX;
    const int j = ty + blockIdx.y * by;
    const int k = (D == 3) ? (int)(tz + blockIdx.z * bz) : 0;
    const bool row_in = j < g.ny && k < g.nz;
    const long long rowbase = j * g.sy + k * g.sz;
    const int tbase = TX * (ty + by * tz);
    for (int s = 0; s < cx; ++s) {
        const int i = ib + s * bx;
        if (row_in && i < g.nx) tile[tbase + tx + s * bx] = gin[rowbase + i];
    }
    __syncthreads();
    if (!interior1(j, g.ny, R)) return;
    if (D == 3 && !interior1(k, g.nz, R)) return;
    const int ZS = TX * by;  // LDS plane stride
    for (int s = 0; s < cx; ++s) {
        const int i = ib + s * bx;
        if (!interior1(i, g.nx, R)) continue;
        const int lx = tx + s * bx;
        const int t = tbase + lx;
        const long long idx = rowbase + i;
        T u = 0;
#pragma unroll
        for (int d = R; d >= 1; --d) u += (lx - d >= 0) ? tile[t - d] : gin[idx - d];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += (lx + d < TX) ? tile[t + d] : gin[idx + d];
#pragma unroll
        for (int d = R; d >= 1; --d) u += (ty - d >= 0) ? tile[t - d * TX] : gin[idx - d * g.sy];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += (ty + d < by) ? tile[t + d * TX] : gin[idx + d * g.sy];
        if constexpr (D == 3) {
#pragma unroll
            for (int d = R; d >= 1; --d) u += (tz - d >= 0) ? tile[t - d * ZS] : gin[idx - d * g.sz];
#pragma unroll
            for (int d = 1; d <= R; ++d) u += (tz + d < bz) ? tile[t + d * ZS] : gin[idx + d * g.sz];
        }
        out[idx] = finish<OP>(u, tile[t], kc);
    }
}

// ---------------------------------------------------------------- SMEM_PADDED

template <int OP, int D, int RS, typename T>
__global__ void k_padded(const T *__restrict__ in, T *__restrict__ out, Geom g, int cx, int pad,
                         T kc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *tile = reinterpret_cast<T *>(lds_raw);
    const int R = RS ? RS : g.r;
    const int bx = blockDim.x, by = blockDim.y, bz = (D == 3) ? blockDim.z : 1;
    const int tx = threadIdx.x, ty = threadIdx.y, tz = (D == 3) ? threadIdx.z : 0;
    const int TX = bx * cx;
    const int W = TX + 2 * R;         // loaded row length
    const int PX = W + pad;           // LDS row pitch
    const int PY = by + 2 * R;
    const int PZ = (D == 3) ? bz + 2 * R : 1;
    const int x0 = blockIdx.x * TX - R;
    const int y0 = blockIdx.y * by - R;
    const int z0 = (D == 3) ? (int)(blockIdx.z * bz) - R : 0;
    const int nthr = bx * by * bz;
    const int tid = tx + bx * (ty + by * tz);
    // Cooperative load: consecutive threads take consecutive x of one row.
    const int rows = PY * PZ;
    for (int row = tid / W, lx = tid % W; row < rows;) {
        const int ly = row % PY, lz = row / PY;
        const int gxx = x0 + lx, gyy = y0 + ly, gzz = z0 + lz;
        if (gxx >= 0 && gxx < g.nx && gyy >= 0 && gyy < g.ny && gzz >= 0 && gzz < g.nz)
            tile[lx + PX * (ly + PY * lz)] = in[gxx + gyy * g.sy + gzz * g.sz];
        lx += nthr;
        while (lx >= W) { lx -= W; ++row; }
    }
    __syncthreads();
    const int j = ty + blockIdx.y * by;
    const int k = (D == 3) ? (int)(tz + blockIdx.z * bz) : 0;
    if (!interior1(j, g.ny, R)) return;
    if (D == 3 && !interior1(k, g.nz, R)) return;
    const int ZS = PX * PY;
    for (int s = 0; s < cx; ++s) {
        const int i = tx + s * bx + blockIdx.x * TX;
        if (!interior1(i, g.nx, R)) continue;
        const int t = (tx + s * bx + R) + PX * ((ty + R) + PY * (D == 3 ? tz + R : 0));
        T u = 0;
#pragma unroll
        for (int d = R; d >= 1; --d) u += tile[t - d];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += tile[t + d];
#pragma unroll
        for (int d = R; d >= 1; --d) u += tile[t - d * PX];
#pragma unroll
        for (int d = 1; d <= R; ++d) u += tile[t + d * PX];
        if constexpr (D == 3) {
#pragma unroll
            for (int d = R; d >= 1; --d) u += tile[t - d * ZS];
#pragma unroll
            for (int d = 1; d <= R; ++d) u += tile[t + d * ZS];
        }
        out[i + j * g.sy + k * g.sz] = finish<OP>(u, tile[t], kc);
    }
}

// ---------------------------------------------------------------- host side

static Geom make_geom(const Cfg &c)
{
    Geom g;
    g.nx = c.nx; g.ny = c.ny; g.nz = c.nz; g.r = c.r;
    g.sy = c.nx; g.sz = (long long)c.nx * c.ny;
    return g;
}

template <typename T> static T kcoef(const Cfg &c)
{
    return c.op == OP_THESIS ? (T)(2 * c.dims * c.r) : (T)c.weight;
}

static st_status finish_grid(Cfg &c)
{
    const long long tx = (long long)c.bx * c.cx;
    c.gx = (int)((c.nx + tx - 1) / tx);
    c.gy = (c.ny + c.by - 1) / c.by;
    c.gz = c.dims == 3 ? (c.nz + c.bz - 1) / c.bz : 1;
    if (c.gy > 65535 || c.gz > 65535)
        return fail(ST_EINVAL, "grid (%d,%d,%d) exceeds the y/z launch limits", c.gx, c.gy, c.gz);
    c.zchunk = 0;
    c.tsteps = 1;
    return ST_OK;
}

// HEURISTIC picks the workgroup shape only; coarsen_x is honoured as given
// (src/stencil_dispatch.cu:14-42 selects the unrolled kernels from
// COARSEN_X whatever the block shape), except that the padded tile halves it
// until the tile fits in LDS.
st_status resolve_base(Cfg &c)
{
    if (c.bx <= 0) { c.bx = 64; c.by = 4; c.bz = 1; }
    c.lds = 0;
    return finish_grid(c);
}

st_status resolve_smem(Cfg &c)
{
    if (c.bx <= 0) { c.bx = 64; c.by = 4; c.bz = c.dims == 3 ? 4 : 1; }
    const int es = c.dtype == ST_F64 ? 8 : 4;
    const long long lds = (long long)c.bx * c.cx * c.by * (c.dims == 3 ? c.bz : 1) * es;
    if (lds > kMaxLds) return fail(ST_ELDS, "SMEM tile needs %lld bytes of LDS", lds);
    c.lds = (int)lds;
    return finish_grid(c);
}

st_status resolve_padded(Cfg &c)
{
    const int es = c.dtype == ST_F64 ? 8 : 4;
    auto tile_lds = [&]() {
        const long long px = (long long)c.bx * c.cx + 2 * c.r + c.pad;
        return px * (c.by + 2 * c.r) * (c.dims == 3 ? c.bz + 2 * c.r : 1) * es;
    };
    if (c.bx <= 0) {
        c.bx = 64; c.by = 4; c.bz = c.dims == 3 ? 4 : 1;
        if (c.by < c.r) c.by = c.r;
        if (c.dims == 3 && c.bz < c.r) c.bz = c.r;
        while ((long long)c.bx * c.by * c.bz > kMaxThreads && c.bx > 1) c.bx /= 2;
        while (tile_lds() > kMaxLds && c.cx > 1) c.cx /= 2;
    }
    const long long lds = tile_lds();
    if (lds > kMaxLds) return fail(ST_ELDS, "SMEM_PADDED tile needs %lld bytes of LDS", lds);
    c.lds = (int)lds;
    return finish_grid(c);
}

template <int OP, int D, int RS, typename T>
static void go(int fam, const Cfg &c, const void *in, void *out, hipStream_t s)
{
    const dim3 grid(c.gx, c.gy, c.gz), block(c.bx, c.by, D == 3 ? c.bz : 1);
    const Geom g = make_geom(c);
    const T *pin = static_cast<const T *>(in);
    T *pout = static_cast<T *>(out);
    if (fam == 0)
        hipLaunchKernelGGL((k_base<OP, D, RS, T>), grid, block, 0, s, pin, pout, g, c.cx, kcoef<T>(c));
    else if (fam == 1)
        hipLaunchKernelGGL((k_smem<OP, D, RS, T>), grid, block, c.lds, s, pin, pout, g, c.cx,
                           kcoef<T>(c));
    else
        hipLaunchKernelGGL((k_padded<OP, D, RS, T>), grid, block, c.lds, s, pin, pout, g, c.cx,
                           c.pad, kcoef<T>(c));
}

template <int OP, int D, typename T>
static void pick_r(int fam, const Cfg &c, const void *in, void *out, hipStream_t s)
{
    switch (c.r) {
    case 1: go<OP, D, 1, T>(fam, c, in, out, s); break;
    case 2: go<OP, D, 2, T>(fam, c, in, out, s); break;
    case 4: go<OP, D, 4, T>(fam, c, in, out, s); break;
    default: go<OP, D, 0, T>(fam, c, in, out, s); break;
    }
}

template <typename T>
static void pick(int fam, const Cfg &c, const void *in, void *out, hipStream_t s)
{
    if (c.op == OP_THESIS) {
        if (c.dims == 3) pick_r<OP_THESIS, 3, T>(fam, c, in, out, s);
        else pick_r<OP_THESIS, 2, T>(fam, c, in, out, s);
    } else {
        if (c.dims == 3) pick_r<OP_JACOBI, 3, T>(fam, c, in, out, s);
        else pick_r<OP_JACOBI, 2, T>(fam, c, in, out, s);
    }
}

static st_status launch_family(int fam, const Cfg &c, const void *in, void *out, hipStream_t s)
{
    if (c.dtype == ST_F64) pick<double>(fam, c, in, out, s);
    else pick<float>(fam, c, in, out, s);
    return hip_check(hipGetLastError(), "sweep launch");
}

st_status launch_base(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    return launch_family(0, c, in, out, s);
}

st_status launch_smem(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    return launch_family(1, c, in, out, s);
}

st_status launch_padded(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    return launch_family(2, c, in, out, s);
}

}  // namespace st
