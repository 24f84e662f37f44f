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
//     problem fits it, else wave tiles of the register-queue march (64
//     columns x TY rows x zchunk planes);
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
    int ty;         // rows per wave (1, 2, 4)
    int tiles_x, tiles_y;
    int zchunk;
    int item0;      // first item of this problem in the flat list
    float *u1, *u2;
};

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
__device__ __forceinline__ void batch_item(const BatchProblem *__restrict__ P, int np, int item,
                                           int sweep)
{
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P[mid].item0 <= item) lo = mid;
        else hi = mid - 1;
    }
    const BatchProblem &q = P[lo];
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

__global__ __launch_bounds__(256) void k_batch_sweep(const BatchProblem *__restrict__ P, int np,
                                                     int nitems, int sweep)
{
    const int item = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (item >= nitems) return;
    batch_item(P, np, item, sweep);
}

__global__ __launch_bounds__(256) void k_batch_persistent(const BatchProblem *__restrict__ P,
                                                          int np, int nitems, int nsweeps)
{
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int w0 = blockIdx.x * wpb + (threadIdx.x >> 6);
    for (int s = 0; s < nsweeps; ++s) {
        for (int item = w0; item < nitems; item += nw) batch_item(P, np, item, s);
        if (s + 1 < nsweeps) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- plane-chunk form
//
// A workgroup of 256 threads owns a tile: frame rows [y0, y0+F) of one
// problem over a z chunk.  A frame's rows of one plane, with R halo rows on
// each side, are one contiguous run of memory, so every plane is fetched
// with fully coalesced loads of consecutive floats (any nx), prefetched
// into registers one plane ahead and written to a double-buffered LDS
// plane (two planes in flight).  Each thread owns up to kBpCells cells of the frame (t, t+256, ...,
// contiguous in memory as well); their x and y neighbours are read from the
// LDS plane and z runs through a ring of R planes and R partial sums, in the
// reference's order, as in sweep_rows.hip.
constexpr int kBpThreads = 256;
constexpr int kBpCells = 8;    // owned cells per thread: frame core <= 2048 cells
constexpr int kBpLoads = 16;   // prefetched floats per thread: frame + halo <= 4096

template <int OP, int R>
__device__ __forceinline__ void bp_tile(const BatchProblem &q, int local, int sweep, float *lds)
{
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
    const int ldoff = (rlo - (y0 - R)) * nx;  // LDS offset of row rlo
    const int plane = (fr + 2 * R) * nx;      // LDS floats per buffer

    // owned cells: interior mask (x, y, plus the z test per plane)
    unsigned own = 0, intr = 0;
#pragma unroll
    for (int i = 0; i < kBpCells; ++i) {
        const int c = t + i * kBpThreads;
        if (c < ncore) {
            own |= 1u << i;
            const int r = c / nx, x = c - r * nx;
            const int y = y0 + r;
            if (x >= R && x < nx - R && y >= R && y < ny - R) intr |= 1u << i;
        }
    }

    float ring[R][kBpCells], part[R][kBpCells], pf[kBpLoads], pf2[kBpLoads];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int i = 0; i < kBpCells; ++i) ring[j][i] = part[j][i] = 0.0f;

    auto load_plane = [&](int k, float (&d)[kBpLoads]) {
        k = min(max(k, 0), nz - 1);
        const float *p = in + (long long)k * sz + (long long)rlo * nx;
#pragma unroll
        for (int i = 0; i < kBpLoads; ++i) {
            const int j = t + i * kBpThreads;
            if (j < nload) d[i] = p[j];
        }
    };
    // two planes in flight: pf holds plane m when step m starts, pf2 plane
    // m+1
    const int kin_lo = z0 - R, kin_hi = z1 + R;
    load_plane(kin_lo, pf);
    if (kin_lo + 1 < kin_hi) load_plane(kin_lo + 1, pf2);
    for (int m = kin_lo; m < kin_hi; ++m) {
        float *buf = lds + (m & 1) * plane;
#pragma unroll
        for (int i = 0; i < kBpLoads; ++i) {
            const int j = t + i * kBpThreads;
            if (j < nload) buf[ldoff + j] = pf[i];
        }
#pragma unroll
        for (int i = 0; i < kBpLoads; ++i) pf[i] = pf2[i];
        if (m + 2 < kin_hi) load_plane(m + 2, pf2);
        __syncthreads();
        const int k = m - R;  // plane completed this step
        const bool kin = k >= z0 && k < z1 && k >= R && k < nz - R;
        float *orow = out + (long long)k * sz + (long long)y0 * nx;
#pragma unroll
        for (int i = 0; i < kBpCells; ++i) {
            if (!((own >> i) & 1u)) continue;
            const int b = R * nx + t + i * kBpThreads;  // LDS index of the cell
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
            if (kin && ((intr >> i) & 1u))
                orow[t + i * kBpThreads] = finish<OP>(part[0][i], ring[0][i], q.kc);
#pragma unroll
            for (int j = 0; j + 1 < R; ++j) {
                ring[j][i] = ring[j + 1][i];
                part[j][i] = part[j + 1][i];
            }
            ring[R - 1][i] = cen;
            part[R - 1][i] = s;
        }
    }
    __syncthreads();  // the next tile rewrites the LDS planes
}

__device__ __forceinline__ void bp_item(const BatchProblem *__restrict__ P, int np, int item, int sweep,
                                        float *lds)
{
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P[mid].item0 <= item) lo = mid;
        else hi = mid - 1;
    }
    const BatchProblem &q = P[lo];
    if (sweep >= q.iters) return;
    const int local = item - q.item0;
    if (q.g.r == 1) {
        if (q.op == OP_THESIS) bp_tile<OP_THESIS, 1>(q, local, sweep, lds);
        else bp_tile<OP_JACOBI, 1>(q, local, sweep, lds);
    } else {
        if (q.op == OP_THESIS) bp_tile<OP_THESIS, 2>(q, local, sweep, lds);
        else bp_tile<OP_JACOBI, 2>(q, local, sweep, lds);
    }
}

__global__ __launch_bounds__(kBpThreads) void k_bp_sweep(const BatchProblem *__restrict__ P, int np, int ntiles,
                                                         int sweep)
{
    __shared__ float lds[2 * kBpLoads * kBpThreads];
    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    if (tile < ntiles) bp_item(P, np, tile, sweep, lds);
}

__global__ __launch_bounds__(kBpThreads) void k_bp_persistent(const BatchProblem *__restrict__ P, int np,
                                                              int ntiles, int nsweeps)
{
    __shared__ float lds[2 * kBpLoads * kBpThreads];
    const int b0 = (int)xcd_remap(blockIdx.x, gridDim.x);
    for (int s = 0; s < nsweeps; ++s) {
        for (int tile = b0; tile < ntiles; tile += (int)gridDim.x) bp_item(P, np, tile, s, lds);
        if (s + 1 < nsweeps) cooperative_groups::this_grid().sync();
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
// plus 2R halo rows fits the LDS plane.
static bool bp_fits(const st_problem &p)
{
    return p.radius >= 1 && p.radius <= 2 && p.nx <= kBpCells * kBpThreads &&
           (1 + 2 * p.radius) * p.nx <= kBpLoads * kBpThreads;
}

// Frame rows: the rule's coarsen_y when given, else as many rows as keep the
// frame within kBpCells cells per thread and the frame plus halo within the
// LDS plane.
static int bp_rows(const st_problem &p, int want)
{
    const int lim = std::min(kBpCells * kBpThreads / p.nx, kBpLoads * kBpThreads / p.nx - 2 * p.radius);
    int f = want > 0 ? std::min(want, lim) : lim;
    if (f > p.ny) f = p.ny;
    return f < 1 ? 1 : f;
}

st_status build_batch(const st_problem *probs, int n, const st_tile_rule *rules, int nrules,
                      std::vector<BatchProblem> &out, int &nitems, int &maxit, bool plane_form)
{
    out.clear();
    nitems = 0;
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
        b.u1 = static_cast<float *>(p.u1);
        b.u2 = static_cast<float *>(p.u2);
        const long long items = (long long)b.tiles_x * b.tiles_y * ((p.nz + zc - 1) / zc);
        if (nitems + items > (1LL << 30)) return fail(ST_EINVAL, "batch too large");
        nitems += (int)items;
        maxit = std::max(maxit, p.iterations);
        out.push_back(b);
    }
    return ST_OK;
}

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
    bool plane_form = true;
    for (int i = 0; i < n; ++i) plane_form = plane_form && bp_fits(probs[i]);
    std::vector<BatchProblem> table;
    int nitems = 0, maxit = 0;
    st_status st = build_batch(probs, n, rules, nrules, table, nitems, maxit, plane_form);
    if (st != ST_OK) return st;
    void *ws = nullptr;
    const size_t bytes = table.size() * sizeof(BatchProblem);
    if ((st = workspace_slot(1, bytes, &ws)) != ST_OK) return st;
    if ((st = hip_check(hipMemcpyAsync(ws, table.data(), bytes, hipMemcpyHostToDevice, s),
                        "batch table copy")) != ST_OK)
        return st;
    if ((st = hip_check(hipStreamSynchronize(s), "batch table sync")) != ST_OK) return st;
    const BatchProblem *P = static_cast<const BatchProblem *>(ws);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
    }
    // work items: wave tiles (4 per workgroup) of the march form, workgroup
    // tiles of the plane-chunk form
    const int groups = plane_form ? nitems : (nitems + 3) / 4;
    if (mode == ST_BATCH_PERSISTENT) {
        const void *fn = plane_form ? reinterpret_cast<const void *>(&k_bp_persistent)
                                    : reinterpret_cast<const void *>(&k_batch_persistent);
        int per_cu = 0, dev = 0, cus = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0);
        if (e != hipSuccess) return hip_check(e, "occupancy");
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int grid = per_cu * cus;
        if (grid > groups) grid = groups;
        if (grid < 1) grid = 1;
        int np = n, ni = nitems, ns = maxit;
        void *args[] = {&P, &np, &ni, &ns};
        if (elapsed_ms) (void)hipEventRecord(e0, s);
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), args, 0, s);
        if (e != hipSuccess) return hip_check(e, "batched cooperative launch");
    } else {
        // capture one launch per sweep, replay as one graph
        const dim3 grid(groups), block(256);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_check(e, "begin capture");
        for (int sw = 0; sw < maxit; ++sw) {
            if (plane_form) hipLaunchKernelGGL(k_bp_sweep, grid, block, 0, s, P, n, nitems, sw);
            else hipLaunchKernelGGL(k_batch_sweep, grid, block, 0, s, P, n, nitems, sw);
        }
        e = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) return hip_check(e, "end capture");
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return hip_check(e, "graph instantiate");
        }
        if (elapsed_ms) (void)hipEventRecord(e0, s);
        e = hipGraphLaunch(exec, s);
        if (e == hipSuccess && elapsed_ms) (void)hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
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
