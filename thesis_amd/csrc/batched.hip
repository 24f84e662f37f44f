// Batched sweeps: many independent grids of uneven sizes advanced together.
//
// No reference counterpart (the reference launches one grid per process,
// src/main.cu); each problem follows the same per-sweep contract as
// stencil_dispatch (ping-pong between its u1 and u2, result in u1 for an
// even iteration count, bit-exact with the CPU oracle).
//
//   * all problems' work is one flat list of tiles ordered by problem; a
//     tile finds its problem by binary search over the list offsets.  The
//     tiles are workgroup tiles of the plane-chunk form (below) when every
//     problem fits it and all share one operator, else wave tiles of the
//     register-queue march (64 columns x TY rows x zchunk planes);
//   * the plane form first compares each problem's two shells; with equal
//     shells it stores whole rows (shell cells with their own value);
//   * ST_BATCH_GRAPH captures one launch per sweep (the whole batch in each)
//     into a hipGraph and replays it as a single graph launch;
//   * ST_BATCH_PERSISTENT runs every sweep inside one cooperative kernel
//     with a grid-wide barrier between sweeps;
//   * ST_BATCH_SEQUENTIAL is the baseline: stencil_dispatch per problem.
// Tile shapes (rows per wave, planes per chunk) come from a rule table that
// the autotuner writes (thesis_amd/tuner); the first rule whose radius
// matches and whose max_cells bounds the problem applies.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "st_device.hpp"
#include "st_internal.hpp"
#include "st_march.hpp"

namespace st {

struct BatchProblem {
    Geom g;
    float kc;       // THESIS divisor or JACOBI weight
    int op;
    int iters;
    int ty;         // rows per wave (1, 2, 4); plane form: frame rows
    int tiles_x, tiles_y;
    int zchunk;
    int item0;      // first item of this problem in the flat list
    // plane form, two fused sweeps per pass (equal shells): frame rows, y
    // tiles, first item in the fused list; n2 fused passes (even) come
    // first, then iters - 2*n2 single sweeps (0: single sweeps only)
    int tyf, tiles_yf, item0f;
    int zchunkf;    // planes per fused tile
    int n2;
    int ncf, nlf;   // fused form: level-1 cells and level-0 load slots per thread
    float *u1, *u2;
};

// Index of the problem that owns a flat-list item: binary search over the
// problems' first items, a compact int array (a few KB for hundreds of
// problems, so the searches hit the scalar cache).
__device__ __forceinline__ int find_problem(const int *__restrict__ starts, int np, int item)
{
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int R, int TY>
__device__ __forceinline__ void batch_tile(const BatchProblem &q, int local, int sweep)
{
    const int tx = local % q.tiles_x;
    const int rest = local / q.tiles_x;
    const float *in = (sweep & 1) ? q.u2 : q.u1;
    float *out = (sweep & 1) ? q.u1 : q.u2;
    if (q.op == OP_THESIS)
        march_tile<OP_THESIS, 3, R, 1, TY, float>(in, out, q.g, tx, rest % q.tiles_y,
                                                   rest / q.tiles_y, q.zchunk, q.kc, 1);
    else
        march_tile<OP_JACOBI, 3, R, 1, TY, float>(in, out, q.g, tx, rest % q.tiles_y,
                                                   rest / q.tiles_y, q.zchunk, q.kc, 1);
}

// Advance the wave's item by one sweep.  Items past a problem's iteration
// count do nothing.
__device__ __forceinline__ void batch_item(const BatchProblem *__restrict__ P, const int *__restrict__ starts,
                                           int np, int item, int sweep)
{
    const BatchProblem &q = P[find_problem(starts, np, item)];
    if (sweep >= q.iters) return;
    const int local = item - q.item0;
    const int r = q.g.r, ty = q.ty;
    if (r == 1) {
        if (ty == 4) batch_tile<1, 4>(q, local, sweep);
        else if (ty == 2) batch_tile<1, 2>(q, local, sweep);
        else batch_tile<1, 1>(q, local, sweep);
    } else {
        if (ty == 2) batch_tile<2, 2>(q, local, sweep);
        else batch_tile<2, 1>(q, local, sweep);
    }
}

// Items [item_lo, item_hi) of the flat list, one wave each.
__global__ __launch_bounds__(256) void k_batch_sweep(const BatchProblem *__restrict__ P,
                                                     const int *__restrict__ starts, int np,
                                                     int item_lo, int item_hi, int sweep)
{
    const int item = item_lo + (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (item >= item_hi) return;
    batch_item(P, starts, np, item, sweep);
}

__global__ __launch_bounds__(256) void k_batch_persistent(const BatchProblem *__restrict__ P,
                                                          const int *__restrict__ starts,
                                                          int np, int nitems, int nsweeps, int *coop_ok)
{
    // a launch that is not cooperative returns before any grid barrier and
    // leaves coop_ok at 0 (the host reports it)
    if (!cooperative_groups::this_grid().is_valid()) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *coop_ok = 1;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int w0 = blockIdx.x * wpb + (threadIdx.x >> 6);
    for (int s = 0; s < nsweeps; ++s) {
        for (int item = w0; item < nitems; item += nw) batch_item(P, starts, np, item, s);
        if (s + 1 < nsweeps) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- plane-chunk form
//
// A workgroup of 512 threads owns a tile: frame rows [y0, y0+F) of one
// problem over a z chunk.  A frame's rows of one plane, with R halo rows on
// each side, are one contiguous run of memory, so every plane is fetched
// with fully coalesced loads of consecutive floats (any nx), prefetched
// into registers two planes ahead and written to a double-buffered LDS
// plane.  Each thread owns up to kBpCells cells of the frame (t, t+512, ...,
// contiguous in memory as well); their x and y neighbours are read from the
// LDS plane and z runs through a ring of R planes and R partial sums, in the
// reference's order, as in sweep_rows.hip.
constexpr int kBpThreads = 512;
constexpr int kBpCells = 4;    // owned cells per thread at R = 1: frame core <= 2048 cells
// owned cells per thread by radius (R = 2 keeps half, for registers)
__host__ __device__ constexpr int bp_cells(int r) { return r == 1 ? kBpCells : kBpCells / 2; }
constexpr int kBpLoads = 8;    // prefetched floats per thread: frame + halo <= 4096
// LDS floats per plane buffer: the last load slot's lanes past the window
// write into the padding
constexpr int kBpStride = kBpLoads * kBpThreads + kBpThreads;
// LDS array: cells past the frame (up to bp_cells(R) * kBpThreads) still
// read their y neighbours, at most 2R * nx + bp_cells(R) * kBpThreads - 1 <
// 4800 floats into a buffer since (1 + 2R) * nx <= 4096 (R = 1: 2730 + 2047,
// R = 2: 3276 + 1023)
constexpr int kBpLds = kBpStride + 4800;

template <int OP, int R>
__device__ __forceinline__ void bp_tile(const BatchProblem &q, int local, int sweep, bool eq, float *lds)
{
    constexpr int NC = bp_cells(R);
    const int t = threadIdx.x;
    const int nx = q.g.nx, ny = q.g.ny, nz = q.g.nz;
    const long long sz = q.g.sz;
    const int ty = local % q.tiles_y, tz = local / q.tiles_y;
    const int y0 = ty * q.ty;                 // ty holds the frame height F
    const int fr = min(q.ty, ny - y0);        // rows of this frame
    const int ncore = fr * nx;
    const int z0 = tz * q.zchunk, z1 = min(z0 + q.zchunk, nz);
    const float *in = (sweep & 1) ? q.u2 : q.u1;
    float *out = (sweep & 1) ? q.u1 : q.u2;
    // contiguous load window of a plane: rows [rlo, rhi) of the domain
    const int rlo = max(y0 - R, 0), rhi = min(y0 + fr + R, ny);
    const int nload = (rhi - rlo) * nx;
    const int nslot = (nload + kBpThreads - 1) / kBpThreads;  // load slots in use
    const int ldoff = (rlo - (y0 - R)) * nx;  // LDS offset of row rlo

    // owned cells: interior mask (x, y, plus the z test per plane), and the
    // mask of stored cells: the interior, or with equal shells whole rows of
    // the interior y range (a shell cell is written with the value it has,
    // so stores cover whole cache lines; partially written lines cost a
    // read-modify-write in HBM)
    unsigned intr = 0, put = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = t + i * kBpThreads;
        if (c < ncore) {
            const int r = c / nx, x = c - r * nx;
            const int y = y0 + r;
            if (x >= R && x < nx - R && y >= R && y < ny - R) intr |= 1u << i;
            if (y >= R && y < ny - R && (eq || (x >= R && x < nx - R))) put |= 1u << i;
        }
    }

    float ring[R][NC], part[R][NC], pf[kBpLoads], pf2[kBpLoads];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int i = 0; i < NC; ++i) ring[j][i] = part[j][i] = 0.0f;

    // Plane loads are raw buffer loads over the input buffer: slots past the
    // window get an offset past the buffer and return 0 without touching
    // memory, lanes of the last slot past the window read the following
    // floats (or 0 past the buffer's end) into the LDS padding.  Every slot
    // is issued and written to LDS, so the compiler's wait counts stay exact.
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc((void *)in, (short)0, (int)(sz * nz * 4), 0x00020000);
    // Results go out through raw buffer stores as well: a cell that is not
    // written gets an offset past the buffer and the store is dropped, so
    // every step issues the same number of stores.
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc((void *)out, (short)0, (int)(sz * nz * 4), 0x00020000);
    // A plane past the chunk (valid false) is issued all past the buffer.
    // Bit 31 of an offset (past any buffer this form takes) marks a load
    // that is off; bit arithmetic rather than selects keeps the loads free
    // of branches.
    unsigned slot_off[kBpLoads];
#pragma unroll
    for (int i = 0; i < kBpLoads; ++i) slot_off[i] = (unsigned)(i >= nslot) << 31;
    auto load_plane = [&](int k, float (&d)[kBpLoads], bool valid) {
        k = min(max(k, 0), nz - 1);
        const unsigned base = ((unsigned)(k * sz + (long long)rlo * nx) * 4u) | ((unsigned)!valid << 31);
#pragma unroll
        for (int i = 0; i < kBpLoads; ++i) {
            const unsigned off = (base | slot_off[i]) + (unsigned)(t + i * kBpThreads) * 4u;
            d[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, (int)off, 0, 0));
        }
    };
    // One step consumes the plane in `pa` (LDS buffer m & 1) and refills
    // `pa` with plane m + 2; the loop alternates pf and pf2, so every load
    // has two steps to land and no register copies force an early wait.
    // Every step issues the same memory instructions (16 loads, NC stores,
    // masked by offset), so the compiler waits for exactly the plane a step
    // consumes.
    const int kin_lo = z0 - R, kin_hi = z1 + R;
    auto step = [&](int m, float (&pa)[kBpLoads]) {
        // keep the compiler from hoisting this step's LDS writes (and so
        // the wait for its plane) into the previous step
        __builtin_amdgcn_sched_barrier(0);
        float *buf = lds + (m & 1) * kBpStride;
        // slots past the window land in the buffer's padding (which holds
        // only such filler), so every load is consumed by one LDS write
#pragma unroll
        for (int i = 0; i < kBpLoads; ++i)
            buf[(i < nslot ? ldoff + i * kBpThreads : kBpLoads * kBpThreads) + t] = pa[i];
        load_plane(m + 2, pa, m + 2 < kin_hi);
        __syncthreads();
        const int k = m - R;  // plane completed this step
        const bool kin = k >= z0 && k < z1 && k >= R && k < nz - R;
        const unsigned obase = (unsigned)(k * sz + (long long)y0 * nx) * 4u;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            // cells past the frame read inside the LDS array and store nothing
            const int b = R * nx + t + i * kBpThreads;
            float s = 0.0f;
#pragma unroll
            for (int d = R; d >= 1; --d) s += buf[b - d];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += buf[b + d];
#pragma unroll
            for (int d = R; d >= 1; --d) s += buf[b - d * nx];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += buf[b + d * nx];
#pragma unroll
            for (int j = 0; j < R; ++j) s += ring[j][i];
            const float cen = buf[b];
#pragma unroll
            for (int j = 0; j < R; ++j) part[j][i] += cen;
            const unsigned off = (kin && ((put >> i) & 1u)) ? obase + (unsigned)(t + i * kBpThreads) * 4u
                                                            : 0x80000000u;
            const float res = ((intr >> i) & 1u) ? finish<OP>(part[0][i], ring[0][i], q.kc) : ring[0][i];
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, res), rout, (int)off, 0, 0);
#pragma unroll
            for (int j = 0; j + 1 < R; ++j) {
                ring[j][i] = ring[j + 1][i];
                part[j][i] = part[j + 1][i];
            }
            ring[R - 1][i] = cen;
            part[R - 1][i] = s;
        }
    };
    // prologue in the loop's instruction pattern (dropped stores between the
    // plane loads), so the wait counts at the loop head match the steady state
    auto store_none = [&]() {
#pragma unroll
        for (int i = 0; i < NC; ++i) __builtin_amdgcn_raw_buffer_store_b32(0u, rout, (int)0x80000000u, 0, 0);
    };
    load_plane(kin_lo, pf, true);
    store_none();
    load_plane(kin_lo + 1, pf2, kin_lo + 1 < kin_hi);
    store_none();
    // an odd step count runs one trailing step on no plane: it loads and
    // stores nothing (k >= z1)
    for (int m = kin_lo; m < kin_hi; m += 2) {
        step(m, pf);
        step(m + 1, pf2);
    }
    __syncthreads();  // the next tile rewrites the LDS planes
}

// ---------------------------------------------------------------- plane form, two fused sweeps
//
// Problems whose two buffers hold the same shell (a shell cell has the same
// value at every time level) advance two sweeps per pass: the workgroup
// streams a frame of F rows over its z chunk, computing level 1 on the
// frame's rows plus R rows on each side from a level-0 window of F + 4R rows,
// and level 2 on the frame's rows from level 1.  Per plane: the level-0
// window goes to LDS (one barrier), level 1 of plane m - R completes in
// registers (ring and partial sums as in the single form) and goes to a
// second LDS plane (one barrier), and level 2 of plane m - 2R completes and
// is stored.  Both LDS planes are single-buffered: the barriers order each
// write after the previous plane's reads.  Each thread owns up to 8 cells of
// each level (flat, consecutive across threads).  A pass reads X = u1 on
// even passes and u2 on odd ones and writes the other; with an even number
// of fused passes per problem the state after them is in u1, as after the
// same number of single sweeps.
//
// Level-1 cells cover the frame's rows plus the halo rows inside the
// domain.  A problem whose whole y extent is one frame takes the smallest
// power of two of cells (and load slots) per thread that covers its plane,
// so small planes do not pay for the largest frame's registers and LDS
// traffic.
// cells per thread and level at most: 8 at R = 1, 4 at R = 2 (registers)
__host__ __device__ constexpr int bp2_cells(int r) { return r == 1 ? 8 : 4; }
constexpr int kBp2L0 = 6144;        // LDS floats of the level-0 window: <= 4096 loaded + slot padding;
                                    // reads of cells past the frame may run on into the level-1 plane
constexpr int kBp2Dummy = kBp2L0 - kBpThreads;  // store target of the unused load slots (past 4096 + 511)
constexpr int kBp2L1 = 7168;        // level-1 plane: level-2 reads reach 3R*nx + 4095 + R <= 6827
constexpr int kBp2Lds = kBp2L0 + kBp2L1;

// Frame rows of the fused form: the level-0 window (F + 4R rows) within the
// 8 load slots of 512 threads and the level-1 rows (F + 2R) within the
// threads' cells; 0 when a single row does not fit.
__host__ __device__ inline int bp2_rows(int nx, int ny, int r)
{
    const int fw = kBpLoads * kBpThreads / nx - 4 * r;
    const int fc = bp2_cells(r) * kBpThreads / nx - 2 * r;
    const int f = fw < fc ? fw : fc;
    return f < 1 ? 0 : (f < ny ? f : ny);
}

// NC level-1 cells per thread, NL level-0 load slots per thread.
template <int OP, int R, int NC, int NL>
__device__ __forceinline__ void bp_tile2(const BatchProblem &q, int local, int pass, float *lds)
{
    static_assert(NC <= bp2_cells(R) && NL <= kBpLoads, "fused tile shape");
    const int t = threadIdx.x;
    const int nx = q.g.nx, ny = q.g.ny, nz = q.g.nz;
    const long long sz = q.g.sz;
    const int ty = local % q.tiles_yf, tz = local / q.tiles_yf;
    const int y0 = ty * q.tyf;
    const int fr = min(q.tyf, ny - y0);
    // level-1 cells: rows l1lo .. l1hi-1, starting c0 cells into the LDS
    // layout whose row 0 is domain row y0 - R.  The smaller shapes (whole-y
    // frames) keep to the rows inside the domain, as their cell count is
    // chosen by; the largest shape takes the frame's rows and R halo rows on
    // each side as they are (one register fewer in its address arithmetic).
    constexpr bool kClamp = NC < bp2_cells(R);
    const int l1lo = kClamp ? max(y0 - R, 0) : y0 - R;
    const int l1hi = kClamp ? min(y0 + fr + R, ny) : y0 + fr + R;
    const int c0 = kClamp ? (l1lo - (y0 - R)) * nx : 0;
    const int n1c = (l1hi - l1lo) * nx;
    const int n2c = fr * nx;            // level-2 cells: rows y0 .. y0+fr-1
    const int z0 = tz * q.zchunkf, z1 = min(z0 + q.zchunkf, nz);
    const float *in = (pass & 1) ? q.u2 : q.u1;
    float *out = (pass & 1) ? q.u1 : q.u2;
    // level-0 window rows [rlo, rhi); LDS row 0 is domain row y0 - 2R
    const int rlo = max(y0 - 2 * R, 0), rhi = min(y0 + fr + 2 * R, ny);
    const int nslot = ((rhi - rlo) * nx + kBpThreads - 1) / kBpThreads;
    const int ldoff = (rlo - (y0 - 2 * R)) * nx;
    float *L0 = lds, *L1 = lds + kBp2L0;

    // x/y interior masks of the owned cells of both levels, and the stored
    // cells of level 2 (whole interior rows: shell cells keep their value)
    unsigned intr1 = 0, intr2 = 0, put = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = t + i * kBpThreads;
        if (c < n1c) {
            const int r = c / nx, x = c - r * nx;
            const int y = l1lo + r;
            if (x >= R && x < nx - R && y >= R && y < ny - R) intr1 |= 1u << i;
        }
        if (c < n2c) {
            const int r = c / nx, x = c - r * nx;
            const int y = y0 + r;
            if (x >= R && x < nx - R && y >= R && y < ny - R) intr2 |= 1u << i;
            if (y >= R && y < ny - R) put |= 1u << i;
        }
    }

    float ring1[R][NC], part1[R][NC], ring2[R][NC], part2[R][NC], pf[NL], pf2[NL];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int i = 0; i < NC; ++i) ring1[j][i] = part1[j][i] = ring2[j][i] = part2[j][i] = 0.0f;

    // raw buffer loads and stores, masked by offset, as in the single form
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc((void *)in, (short)0, (int)(sz * nz * 4), 0x00020000);
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc((void *)out, (short)0, (int)(sz * nz * 4), 0x00020000);
    unsigned slot_off[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) slot_off[i] = (unsigned)(i >= nslot) << 31;
    auto load_plane = [&](int k, float (&d)[NL], bool valid) {
        k = min(max(k, 0), nz - 1);
        const unsigned base = ((unsigned)(k * sz + (long long)rlo * nx) * 4u) | ((unsigned)!valid << 31);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const unsigned off = (base | slot_off[i]) + (unsigned)(t + i * kBpThreads) * 4u;
            d[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, (int)off, 0, 0));
        }
    };
    const int kin_lo = z0 - 2 * R, kin_hi = z1 + 2 * R;
    auto step = [&](int m, float (&pa)[NL]) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) L0[(i < nslot ? ldoff + i * kBpThreads : kBp2Dummy) + t] = pa[i];
        load_plane(m + 2, pa, m + 2 < kin_hi);
        __syncthreads();
        // level 1: plane m - R completes
        const bool z1in = m - R >= R && m - R < nz - R;
        float v1[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const int b = R * nx + c0 + t + i * kBpThreads;
            float s = 0.0f;
#pragma unroll
            for (int d = R; d >= 1; --d) s += L0[b - d];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += L0[b + d];
#pragma unroll
            for (int d = R; d >= 1; --d) s += L0[b - d * nx];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += L0[b + d * nx];
#pragma unroll
            for (int j = 0; j < R; ++j) s += ring1[j][i];
            const float cen = L0[b];
#pragma unroll
            for (int j = 0; j < R; ++j) part1[j][i] += cen;
            v1[i] = (z1in && ((intr1 >> i) & 1u)) ? finish<OP>(part1[0][i], ring1[0][i], q.kc) : ring1[0][i];
#pragma unroll
            for (int j = 0; j + 1 < R; ++j) {
                ring1[j][i] = ring1[j + 1][i];
                part1[j][i] = part1[j + 1][i];
            }
            ring1[R - 1][i] = cen;
            part1[R - 1][i] = s;
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) L1[R * nx + c0 + t + i * kBpThreads] = v1[i];
        __syncthreads();
        // level 2: plane m - 2R completes
        const int k = m - 2 * R;
        const bool kst = k >= z0 && k < z1 && k >= R && k < nz - R;
        const unsigned obase = (unsigned)(k * sz + (long long)y0 * nx) * 4u;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const int b = 2 * R * nx + t + i * kBpThreads;
            float s = 0.0f;
#pragma unroll
            for (int d = R; d >= 1; --d) s += L1[b - d];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += L1[b + d];
#pragma unroll
            for (int d = R; d >= 1; --d) s += L1[b - d * nx];
#pragma unroll
            for (int d = 1; d <= R; ++d) s += L1[b + d * nx];
#pragma unroll
            for (int j = 0; j < R; ++j) s += ring2[j][i];
            const float cen = L1[b];
#pragma unroll
            for (int j = 0; j < R; ++j) part2[j][i] += cen;
            const unsigned off = (kst && ((put >> i) & 1u)) ? obase + (unsigned)(t + i * kBpThreads) * 4u
                                                             : 0x80000000u;
            const float res = ((intr2 >> i) & 1u) ? finish<OP>(part2[0][i], ring2[0][i], q.kc) : ring2[0][i];
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, res), rout, (int)off, 0, 0);
#pragma unroll
            for (int j = 0; j + 1 < R; ++j) {
                ring2[j][i] = ring2[j + 1][i];
                part2[j][i] = part2[j + 1][i];
            }
            ring2[R - 1][i] = cen;
            part2[R - 1][i] = s;
        }
    };
    auto store_none = [&]() {
#pragma unroll
        for (int i = 0; i < NC; ++i) __builtin_amdgcn_raw_buffer_store_b32(0u, rout, (int)0x80000000u, 0, 0);
    };
    load_plane(kin_lo, pf, true);
    store_none();
    load_plane(kin_lo + 1, pf2, kin_lo + 1 < kin_hi);
    store_none();
    for (int m = kin_lo; m < kin_hi; m += 2) {
        step(m, pf);
        step(m + 1, pf2);
    }
    __syncthreads();  // the next tile rewrites the LDS planes
}

// Shell comparison of every problem: flags[p] is set to 1 by k_bp_flags,
// then cleared by any block of k_bp_shell_eq (grid: problems x parts) that
// finds a cell of problem p's R-thick shell whose bits differ between u1 and
// u2.
__global__ __launch_bounds__(256) void k_bp_flags(int *__restrict__ flags, int np, int nzero)
{
    // then nzero words of 0: the persistent launch's cooperative flag and
    // its per-pass tile counters
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < np) flags[i] = 1;
    else if (i < np + nzero) flags[i] = 0;
}

__global__ __launch_bounds__(256) void k_bp_shell_eq(const BatchProblem *__restrict__ P, int *__restrict__ flags)
{
    const BatchProblem &q = P[blockIdx.x];
    const int nx = q.g.nx, ny = q.g.ny, nz = q.g.nz, R = q.g.r;
    const unsigned *a = reinterpret_cast<const unsigned *>(q.u1);
    const unsigned *b = reinterpret_cast<const unsigned *>(q.u2);
    // (a problem of the plane form holds under 2^29 cells)
    const unsigned n1 = (unsigned)shell_full_cells(nx, ny, nz, R), n = (unsigned)shell_cells(nx, ny, nz, R);
    int ok = 1;
    for (unsigned i = blockIdx.y * 256u + threadIdx.x; i < n; i += 256u * gridDim.y) {
        const long long idx = shell_cell(i, n1, nx, ny, nz, R);
        ok &= a[idx] == b[idx];
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0 && !ok) flags[blockIdx.x] = 0;
}

// The plane form is instantiated per operator (a batch takes it when all its
// problems share one), so each kernel inlines only the two radius bodies.
// `pass` counts launches: a problem with n2 fused passes takes its single
// sweeps from pass n2 on, sweep index pass + n2 (the state after 2*n2
// sweeps is in u1, since n2 is even).
template <int OP>
__device__ __forceinline__ void bp_item(const BatchProblem *__restrict__ P, const int *__restrict__ starts,
                                        const int *__restrict__ eqf, int np, int item, int pass, float *lds)
{
    const int pi = find_problem(starts, np, item);
    const BatchProblem &q = P[pi];
    if (pass < q.n2) return;
    const int sweep = pass + q.n2;
    if (sweep >= q.iters) return;
    const int local = item - q.item0;
    const bool eq = eqf[pi] != 0;
    if (q.g.r == 1) bp_tile<OP, 1>(q, local, sweep, eq, lds);
    else bp_tile<OP, 2>(q, local, sweep, eq, lds);
}

// Tiles [tile_lo, tile_hi) of the flat list, one workgroup each.
template <int OP>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(OP == OP_THESIS ? 4 : 6))) void k_bp_sweep(
    const BatchProblem *__restrict__ P, const int *__restrict__ starts, const int *__restrict__ eqf, int np,
    int tile_lo, int tile_hi, int sweep)
{
    __shared__ float lds[kBpLds];
    const int tile = tile_lo + (int)xcd_remap(blockIdx.x, gridDim.x);
    if (tile < tile_hi) bp_item<OP>(P, starts, eqf, np, tile, sweep, lds);
}

// Fused tile `local` of problem q in its tile shape.
template <int OP>
__device__ __forceinline__ void bp_fused_tile(const BatchProblem &q, int local, int pass, float *lds)
{
    if (q.g.r == 1) {
        if (q.ncf == 8) bp_tile2<OP, 1, 8, 8>(q, local, pass, lds);
        else if (q.ncf == 4) bp_tile2<OP, 1, 4, 4>(q, local, pass, lds);
        else if (q.ncf == 2) bp_tile2<OP, 1, 2, 2>(q, local, pass, lds);
        else bp_tile2<OP, 1, 1, 1>(q, local, pass, lds);
    } else {
        // (no 4-cell, 4-slot shape at R = 2: with it the kernel's register
        // allocation spills)
        if (q.ncf == 4) bp_tile2<OP, 2, 4, 8>(q, local, pass, lds);
        else if (q.ncf == 2) bp_tile2<OP, 2, 2, 2>(q, local, pass, lds);
        else bp_tile2<OP, 2, 1, 1>(q, local, pass, lds);
    }
}

// Fused items [tile_lo, tile_hi) of pass `pass`, one workgroup each: problems
// with pass < n2 advance two sweeps.  (THESIS's division needs more
// registers: two waves per SIMD, one workgroup per CU.)
template <int OP>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(OP == OP_THESIS ? 2 : 4))) void k_bp_fused(
    const BatchProblem *__restrict__ P, const int *__restrict__ starts, int np, int tile_lo, int tile_hi, int pass)
{
    __shared__ float lds[kBp2Lds];
    const int tile = tile_lo + (int)xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= tile_hi) return;
    const BatchProblem &q = P[find_problem(starts, np, tile)];
    if (pass >= q.n2) return;
    bp_fused_tile<OP>(q, tile - q.item0f, pass, lds);
}

// Persistent form: one cooperative launch runs every pass of the batch, a
// grid barrier between passes; in pass p each problem takes a fused step
// of two sweeps (p < n2) or its single sweep p + n2 (as one launch per
// pass does in the graph mode).  Workgroups take the pass's tiles (fused
// ones first, the larger) from a counter per pass, as the hardware
// dispatcher hands out a launch's workgroups; the counters start at 0
// (k_bp_flags clears them before every launch).  passmask[p] names the
// lists pass p runs (bit 0 fused, bit 1 single), so that a pass skips the
// list whose tiles would all return at once.  The launch is captured
// into a hipGraph and replayed; a launch that is not cooperative returns
// before any barrier and leaves coop_ok at 0.
template <int OP>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(OP == OP_THESIS ? 2 : 4))) void k_bp_persistent(
    const BatchProblem *__restrict__ P, const int *__restrict__ starts, const int *__restrict__ startsf,
    const int *__restrict__ eqf, const int *__restrict__ passmask, int np, int ntiles, int ntilesf,
    int npasses, int *coop_ok, int *__restrict__ next)
{
    __shared__ float lds[kBp2Lds > kBpLds ? kBp2Lds : kBpLds];
    __shared__ int s_tile;
    if (!cooperative_groups::this_grid().is_valid()) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *coop_ok = 1;
    for (int p = 0; p < npasses; ++p) {
        const int m = passmask[p];
        const int nf = (m & 1) ? ntilesf : 0;
        const int total = nf + ((m & 2) ? ntiles : 0);
        // the first tile of every workgroup is its own index, the rest come
        // from the counter
        for (int tile = (int)blockIdx.x;;) {
            if (tile >= total) break;
            if (tile < nf) {
                const BatchProblem &q = P[find_problem(startsf, np, tile)];
                if (p < q.n2) bp_fused_tile<OP>(q, tile - q.item0f, p, lds);
            } else {
                bp_item<OP>(P, starts, eqf, np, tile - nf, p, lds);
            }
            if (threadIdx.x == 0) s_tile = (int)gridDim.x + atomicAdd(&next[p], 1);
            __syncthreads();
            tile = s_tile;
            __syncthreads();  // read by every thread before the next grab
        }
        if (p + 1 < npasses) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- host side

static const st_tile_rule *match_rule(const st_problem &p, const st_tile_rule *rules, int nrules)
{
    const long long cells = (long long)p.nx * p.ny * p.nz;
    for (int i = 0; i < nrules; ++i)
        if (rules[i].radius == p.radius && cells <= rules[i].max_cells) return &rules[i];
    return nullptr;
}

static int valid_ty(int r, int ty)
{
    if (r == 1) return ty >= 4 ? 4 : (ty >= 2 ? 2 : 1);
    return ty >= 2 ? 2 : 1;
}

// The plane-chunk form takes a batch when every problem's frame of one row
// plus 2R halo rows fits the LDS plane and each buffer is addressable by a
// 32-bit buffer offset.
static bool bp_fits(const st_problem &p)
{
    return p.radius >= 1 && p.radius <= 2 && p.nx <= bp_cells(p.radius) * kBpThreads &&
           (1 + 2 * p.radius) * p.nx <= kBpLoads * kBpThreads &&
           (long long)p.nx * p.ny * p.nz * 4 <= 0x7fffffffLL;
}

// Frame rows: the rule's coarsen_y when given, else as many rows as keep the
// frame within kBpCells cells per thread and the frame plus halo within the
// LDS plane.
static int bp_rows(const st_problem &p, int want)
{
    const int lim = std::min(bp_cells(p.radius) * kBpThreads / p.nx, kBpLoads * kBpThreads / p.nx - 2 * p.radius);
    int f = want > 0 ? std::min(want, lim) : lim;
    if (f > p.ny) f = p.ny;
    return f < 1 ? 1 : f;
}

// Planes per tile of the fused form: its chunks carry 4R halo planes (2R in
// the single form), so they are deeper than the tuned single-form chunks
// (mostly 16 planes).  Measured on the 512-grid batch (50 sweeps, graph
// mode), with one tile shape for all: the tuned chunks 7.09 ms, whole grids
// in z 7.38, 64 planes 6.64, 32 planes 6.49; with per-problem tile shapes
// (three runs each on one box): 32 planes 6.02-6.20, 48 for R = 2 only
// 5.95-6.01, 48 for both radii 5.95 (48 for R = 1 only 6.37-6.40, 40 for
// both 6.15, 64 for R = 1 6.83).
static int fused_chunk(const st_problem &p)
{
    const int want = 48;
    return want > p.nz ? p.nz : want;
}

// Cells and load slots per thread of the fused form: a problem whose whole
// y extent is one frame holds its plane (level 1 and the level-0 window are
// both ny * nx cells) in the smallest power of two per thread below the
// largest shape that covers it; framed problems, and planes that need it,
// take the largest shape (bp2_cells(R) cells, kBpLoads slots).
static void fused_shape(const st_problem &p, int tiles_yf, int &nc, int &nl)
{
    const int top = bp2_cells(p.radius);
    nc = top;
    nl = kBpLoads;
    if (tiles_yf != 1) return;
    const long long plane = (long long)p.nx * p.ny;
    for (int c = 1; c < top; c *= 2)
        if (plane <= (long long)c * kBpThreads) {
            nc = nl = c;
            return;
        }
}

st_status build_batch(const st_problem *probs, int n, const st_tile_rule *rules, int nrules,
                      std::vector<BatchProblem> &out, int &nitems, int &nitemsf_out, int &maxit, bool plane_form)
{
    out.clear();
    nitems = 0;
    nitemsf_out = 0;
    long long nitemsf = 0;
    maxit = 0;
    for (int i = 0; i < n; ++i) {
        const st_problem &p = probs[i];
        st_params sp;
        stencil_params_default(&sp);
        sp.nx = p.nx; sp.ny = p.ny; sp.nz = p.nz; sp.dimensions = 3; sp.radius = p.radius;
        sp.iterations = p.iterations; sp.op = p.op; sp.weight = p.weight; sp.heuristic = 1;
        sp.variant = ST_V_BASE;
        st_status s = stencil_validate(&sp);
        if (s != ST_OK) return fail(s, "problem %d: %s", i, stencil_last_error());
        if (p.radius > 2) return fail(ST_EUNSUPPORTED, "batched sweeps are built for radius 1-2 (problem %d)", i);
        if (!p.u1 || !p.u2 || p.u1 == p.u2) return fail(ST_EINVAL, "problem %d needs two buffers", i);
        const st_tile_rule *rule = match_rule(p, rules, nrules);
        int ty, zc;
        if (plane_form) {
            ty = bp_rows(p, rule ? rule->coarsen_y : 0);
            zc = rule && rule->zchunk > 0 ? rule->zchunk : (p.nz <= 96 ? p.nz : 64);
        } else {
            ty = valid_ty(p.radius, rule ? rule->coarsen_y : 2);
            zc = rule ? rule->zchunk : 32;
            if (zc <= 0) zc = p.nz;
        }
        BatchProblem b;
        std::memset(&b, 0, sizeof b);  // padding too: the table's bytes key the graph cache
        b.g.nx = p.nx; b.g.ny = p.ny; b.g.nz = p.nz; b.g.r = p.radius;
        b.g.sy = p.nx; b.g.sz = (long long)p.nx * p.ny;
        b.op = p.op;
        const double w = p.weight > 0 ? p.weight : 1.0 / (6.0 * p.radius);
        b.kc = p.op == ST_OP_THESIS ? (float)(6 * p.radius) : (float)w;
        b.iters = p.iterations;
        b.ty = ty;
        b.tiles_x = plane_form ? 1 : (p.nx + 63) / 64;
        b.tiles_y = (p.ny + ty - 1) / ty;
        b.zchunk = zc;
        b.item0 = nitems;
        // fused form (chosen per problem once its shells are compared):
        // frames as tall as the level-0 window allows
        b.tyf = plane_form ? bp2_rows(p.nx, p.ny, p.radius) : 0;
        b.tiles_yf = b.tyf > 0 ? (p.ny + b.tyf - 1) / b.tyf : 0;
        fused_shape(p, b.tiles_yf, b.ncf, b.nlf);
        b.item0f = nitemsf;
        b.zchunkf = fused_chunk(p);
        b.n2 = 0;
        nitemsf += (long long)b.tiles_yf * ((p.nz + b.zchunkf - 1) / b.zchunkf);
        b.u1 = static_cast<float *>(p.u1);
        b.u2 = static_cast<float *>(p.u2);
        const long long items = (long long)b.tiles_x * b.tiles_y * ((p.nz + zc - 1) / zc);
        if (nitems + items > (1LL << 30)) return fail(ST_EINVAL, "batch too large");
        nitems += (int)items;
        if (nitemsf > (1LL << 30)) return fail(ST_EINVAL, "batch too large");
        maxit = std::max(maxit, p.iterations);
        out.push_back(b);
    }
    nitemsf_out = (int)nitemsf;
    return ST_OK;
}

// The instantiated graph of the last batch dispatched in graph mode on each
// (device, stream): a dispatch whose table, workspace and launch shape are
// the same bytes replays it instead of capturing and instantiating again.
struct BatchGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
};
static std::mutex g_graph_mutex;
static std::map<std::pair<int, hipStream_t>, BatchGraph> g_graphs;

st_status run_batched(const st_problem *probs, int n, const st_tile_rule *rules, int nrules,
                      int mode, hipStream_t s, float *elapsed_ms)
{
    if (n <= 0 || !probs) return fail(ST_EINVAL, "empty batch");
    if (mode == ST_BATCH_SEQUENTIAL) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (elapsed_ms) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, s);
        }
        for (int i = 0; i < n; ++i) {
            st_params sp;
            stencil_params_default(&sp);
            sp.nx = probs[i].nx; sp.ny = probs[i].ny; sp.nz = probs[i].nz; sp.dimensions = 3;
            sp.radius = probs[i].radius; sp.iterations = probs[i].iterations; sp.op = probs[i].op;
            sp.weight = probs[i].weight; sp.heuristic = 1; sp.variant = ST_V_AUTO;
            int in2 = 0;
            st_status st = stencil_dispatch(&sp, probs[i].u1, probs[i].u2, s, &in2, nullptr);
            if (st != ST_OK) return st;
        }
        if (elapsed_ms) {
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(elapsed_ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return ST_OK;
    }
    {
        // the graph and persistent modes upload the problem table from host
        // memory and read the shell comparison back, then capture and replay
        // their own graph: none of that can be recorded into a caller's
        // capture
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s != nullptr && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return fail(ST_EINVAL, "batched dispatch on a stream under capture (it captures its own graph)");
    }
    bool plane_form = true;
    for (int i = 0; i < n; ++i) plane_form = plane_form && bp_fits(probs[i]) && probs[i].op == probs[0].op;
    const bool thesis = probs[0].op == ST_OP_THESIS;
    std::vector<BatchProblem> table;
    int nitems = 0, nitemsf = 0, maxit = 0;
    st_status st = build_batch(probs, n, rules, nrules, table, nitems, nitemsf, maxit, plane_form);
    if (st != ST_OK) return st;
    // device copy: the problem table, each problem's first item in the
    // single and in the fused list, then the shell-comparison flags
    void *ws = nullptr;
    const size_t tbytes = (table.size() * sizeof(BatchProblem) + 255) & ~(size_t)255;
    const size_t sbytes = (table.size() * sizeof(int) + 255) & ~(size_t)255;
    // shell flags, then the persistent launch's cooperative flag and its tile
    // counter per pass (at most maxit passes)
    const size_t fbytes = (table.size() + 1 + (size_t)maxit) * sizeof(int);
    // then the lists each pass runs (bit 0: fused tiles, bit 1: single-sweep
    // tiles), at most maxit passes
    const size_t mbytes = ((size_t)maxit * sizeof(int) + 255) & ~(size_t)255;
    std::vector<unsigned char> host(tbytes + 2 * sbytes + mbytes, 0);
    auto pack = [&]() {
        std::memcpy(host.data(), table.data(), table.size() * sizeof(BatchProblem));
        for (size_t i = 0; i < table.size(); ++i) {
            std::memcpy(host.data() + tbytes + i * sizeof(int), &table[i].item0, sizeof(int));
            std::memcpy(host.data() + tbytes + sbytes + i * sizeof(int), &table[i].item0f, sizeof(int));
        }
        for (int p = 0; p < maxit; ++p) {
            int m = 0;
            for (const BatchProblem &b : table) {
                if (p < b.n2) m |= 1;
                if (p >= b.n2 && p < b.iters - b.n2) m |= 2;
            }
            std::memcpy(host.data() + tbytes + 2 * sbytes + p * sizeof(int), &m, sizeof(int));
        }
    };
    pack();
    if ((st = workspace_slot(1, host.size() + fbytes, s, &ws)) != ST_OK) return st;
    if ((st = hip_check(hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, s),
                        "batch table copy")) != ST_OK)
        return st;
    if ((st = hip_check(hipStreamSynchronize(s), "batch table sync")) != ST_OK) return st;
    const BatchProblem *P = static_cast<const BatchProblem *>(ws);
    const int *starts = reinterpret_cast<const int *>(static_cast<unsigned char *>(ws) + tbytes);
    const int *startsf = reinterpret_cast<const int *>(static_cast<unsigned char *>(ws) + tbytes + sbytes);
    const int *passmask = reinterpret_cast<const int *>(static_cast<unsigned char *>(ws) + tbytes + 2 * sbytes);
    int *eqf = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + tbytes + 2 * sbytes + mbytes);
    const int *ceqf = eqf;
    int *coop_ok = eqf + n;
    int *next = eqf + n + 1;
    // plane form: compare the problems' shells first (on the stream, inside
    // the timed region)
    // parts per problem: about 8192 shell cells per block of the largest shell
    long long max_shell = 1;
    for (int i = 0; i < n; ++i)
        max_shell = std::max(max_shell, shell_cells(probs[i].nx, probs[i].ny, probs[i].nz, probs[i].radius));
    const int parts = (int)std::min<long long>(256, (max_shell + 8191) / 8192);
    auto shell_check = [&]() {
        hipLaunchKernelGGL(k_bp_flags, dim3((n + 1 + maxit + 255) / 256), dim3(256), 0, s, eqf, n, 1 + maxit);
        if (plane_form) hipLaunchKernelGGL(k_bp_shell_eq, dim3(n, parts), dim3(256), 0, s, P, eqf);
    };
    // Equal shells: fused passes.  The flags are read back, and every problem
    // whose two shells are equal takes n2 = (iters - iters mod 4)/2 fused
    // passes of two sweeps (an even number, so the state after them is in
    // u1 as after the same number of single sweeps), then iters mod 4 single
    // sweeps; the others single sweeps only.  The table goes up again.
    auto fuse_plan = [&](bool &any_fused) -> st_status {
        any_fused = false;
        if (!plane_form) return ST_OK;
        std::vector<int> flags(n, 0);
        hipError_t e = hipMemcpyAsync(flags.data(), eqf, n * sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_check(e, "batch shell flags");
        for (int i = 0; i < n; ++i) {
            BatchProblem &b = table[i];
            b.n2 = (flags[i] && b.tyf > 0) ? (b.iters - b.iters % 4) / 2 : 0;
            any_fused = any_fused || b.n2 > 0;
        }
        if (any_fused) {
            pack();
            e = hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return hip_check(e, "batch table update");
        }
        return ST_OK;
    };
    // passes: problem i runs n2 fused passes, then iters - 2*n2 single
    // sweeps (pass index p >= n2 is its sweep p + n2)
    auto pass_count = [&]() {
        int passes = 0;
        for (const BatchProblem &b : table) passes = std::max(passes, b.iters - b.n2);
        return passes;
    };
    // the graph cache key: the table, the workspace, the launch shape and mode
    auto graph_key = [&]() {
        std::vector<unsigned char> key(host.size() + sizeof(void *) + 8 * sizeof(int));
        unsigned char *k = key.data();
        std::memcpy(k, host.data(), host.size());
        k += host.size();
        std::memcpy(k, &ws, sizeof(void *));
        k += sizeof(void *);
        const int groups_ = plane_form ? nitems : (nitems + 3) / 4;
        const int meta[8] = {n, nitems, nitemsf, maxit, (int)plane_form, (int)thesis, groups_, mode};
        std::memcpy(k, meta, sizeof meta);
        return key;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
    }
    // work items: wave tiles (4 per workgroup) of the march form, workgroup
    // tiles of the plane-chunk form
    const int groups = plane_form ? nitems : (nitems + 3) / 4;
    if (mode == ST_BATCH_PERSISTENT) {
        // One persistent cooperative launch (fused passes where the shells
        // are equal, a grid barrier between passes), captured into a
        // hipGraph that is kept for the next dispatch with the same table,
        // workspace and shape, and replayed.  The null stream takes the
        // cooperative launch directly.
        const void *fn = !plane_form ? reinterpret_cast<const void *>(&k_batch_persistent)
                         : thesis    ? reinterpret_cast<const void *>(&k_bp_persistent<OP_THESIS>)
                                     : reinterpret_cast<const void *>(&k_bp_persistent<OP_JACOBI>);
        int per_cu = 0, dev = 0, cus = 0;
        const int threads = plane_form ? kBpThreads : 256;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0);
        if (e != hipSuccess) return hip_check(e, "occupancy");
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (elapsed_ms) (void)hipEventRecord(e0, s);
        shell_check();
        bool any_fused = false;
        if ((st = fuse_plan(any_fused)) != ST_OK) return st;
        int grid = per_cu * cus;
        grid = std::min(grid, std::max(groups, plane_form ? nitemsf : 0));
        if (grid < 1) grid = 1;
        int np = n, ni = nitems, nif = nitemsf, npass = plane_form ? pass_count() : maxit;
        void *args_bp[] = {&P, &starts, &startsf, &ceqf, &passmask, &np, &ni, &nif, &npass, &coop_ok, &next};
        void *args_march[] = {&P, &starts, &np, &ni, &npass, &coop_ok};
        void **args = plane_form ? args_bp : args_march;
        const bool can_capture = s != nullptr;
        std::vector<unsigned char> key = graph_key();
        std::lock_guard<std::mutex> lock(g_graph_mutex);
        BatchGraph &cached = g_graphs[std::make_pair(dev, s)];
        if (can_capture && (!cached.exec || cached.key != key)) {
            if (cached.exec) {
                (void)hipGraphExecDestroy(cached.exec);
                cached.exec = nullptr;
                cached.key.clear();
            }
            hipGraph_t graph = nullptr;
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_check(e, "begin capture");
            e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, 0, s);
            hipError_t ee = hipStreamEndCapture(s, &graph);
            if (e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return hip_check(e, "batched cooperative launch (capture)");
            }
            if (ee != hipSuccess) return hip_check(ee, "end capture");
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return hip_check(e, "graph instantiate");
            cached.exec = exec;
            cached.key = std::move(key);
        }
        e = can_capture ? hipGraphLaunch(cached.exec, s)
                        : hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, 0, s);
        if (e != hipSuccess) return hip_check(e, "batched persistent launch");
        if (elapsed_ms) (void)hipEventRecord(e1, s);
        int ok = 0;
        e = hipMemcpyAsync(&ok, coop_ok, sizeof ok, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_check(e, "batched persistent launch");
        if (ok != 1) return fail(ST_ELAUNCH, "the persistent batch launch ran without a cooperative grid");
        if (elapsed_ms) {
            (void)hipEventElapsedTime(elapsed_ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return ST_OK;
    } else {
        // Graph mode.  The shells are compared first and the flags read back,
        // so that every problem with equal shells takes fused passes of two
        // sweeps (n2 of them, an even number, then iters mod 4 single
        // sweeps) and the rest single sweeps; then one launch per pass (a
        // fused and/or a single-sweep launch) is captured and replayed as one
        // graph.  The instantiated graph is kept for the next dispatch with
        // the same table (n2 included), workspace and launch shape.
        if (elapsed_ms) (void)hipEventRecord(e0, s);
        bool any_fused = false;
        if (plane_form) {
            shell_check();
            if ((st = fuse_plan(any_fused)) != ST_OK) return st;
        }
        const dim3 block(plane_form ? kBpThreads : 256);
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::vector<unsigned char> key = graph_key();
        std::lock_guard<std::mutex> lock(g_graph_mutex);
        BatchGraph &cached = g_graphs[std::make_pair(dev, s)];
        hipError_t e = hipSuccess;
        if (!cached.exec || cached.key != key) {
            if (cached.exec) {
                (void)hipGraphExecDestroy(cached.exec);
                cached.exec = nullptr;
                cached.key.clear();
            }
            hipGraph_t graph = nullptr;
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_check(e, "begin capture");
            const int passes = pass_count();
            for (int p = 0; p < passes; ++p) {
                bool fused = false, single = false;
                for (const BatchProblem &b : table) {
                    fused = fused || p < b.n2;
                    single = single || (p >= b.n2 && p < b.iters - b.n2);
                }
                if (fused && nitemsf > 0) {
                    const dim3 gf(nitemsf);
                    if (thesis)
                        hipLaunchKernelGGL(k_bp_fused<OP_THESIS>, gf, block, 0, s, P, startsf, n, 0, nitemsf, p);
                    else
                        hipLaunchKernelGGL(k_bp_fused<OP_JACOBI>, gf, block, 0, s, P, startsf, n, 0, nitemsf, p);
                }
                if (single && nitems > 0) {
                    const dim3 gg(plane_form ? nitems : (nitems + 3) / 4);
                    if (plane_form && thesis)
                        hipLaunchKernelGGL(k_bp_sweep<OP_THESIS>, gg, block, 0, s, P, starts, ceqf, n, 0, nitems, p);
                    else if (plane_form)
                        hipLaunchKernelGGL(k_bp_sweep<OP_JACOBI>, gg, block, 0, s, P, starts, ceqf, n, 0, nitems, p);
                    else hipLaunchKernelGGL(k_batch_sweep, gg, block, 0, s, P, starts, n, 0, nitems, p);
                }
            }
            e = hipStreamEndCapture(s, &graph);
            if (e != hipSuccess) return hip_check(e, "end capture");
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return hip_check(e, "graph instantiate");
            cached.exec = exec;
            cached.key = std::move(key);
        }
        e = hipGraphLaunch(cached.exec, s);
        if (e == hipSuccess && elapsed_ms) (void)hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_check(e, "graph launch");
        if (elapsed_ms) {
            (void)hipEventElapsedTime(elapsed_ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return ST_OK;
    }
    if (elapsed_ms) {
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(elapsed_ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return hip_check(hipGetLastError(), "batched launch");
}

}  // namespace st
