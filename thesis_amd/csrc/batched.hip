// Batched sweeps: many independent grids of uneven sizes advanced together.
//
// No reference counterpart (the reference launches one grid per process,
// src/main.cu); each problem follows the same per-sweep contract as
// stencil_dispatch (ping-pong between its u1 and u2, result in u1 for an
// even iteration count, bit-exact with the CPU oracle).
//
//   * all problems' work is one flat list of wave tiles (register-queue
//     march tiles, 64 columns x TY rows x zchunk planes) ordered by problem;
//     a wave finds its problem by binary search over the list offsets;
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

st_status build_batch(const st_problem *probs, int n, const st_tile_rule *rules, int nrules,
                      std::vector<BatchProblem> &out, int &nitems, int &maxit)
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
        int ty = rule ? rule->coarsen_y : 2;
        int zc = rule ? rule->zchunk : 32;
        ty = valid_ty(p.radius, ty);
        if (zc <= 0) zc = p.nz;
        BatchProblem b;
        b.g.nx = p.nx; b.g.ny = p.ny; b.g.nz = p.nz; b.g.r = p.radius;
        b.g.sy = p.nx; b.g.sz = (long long)p.nx * p.ny;
        b.op = p.op;
        const double w = p.weight > 0 ? p.weight : 1.0 / (6.0 * p.radius);
        b.kc = p.op == ST_OP_THESIS ? (float)(6 * p.radius) : (float)w;
        b.iters = p.iterations;
        b.ty = ty;
        b.tiles_x = (p.nx + 63) / 64;
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
    std::vector<BatchProblem> table;
    int nitems = 0, maxit = 0;
    st_status st = build_batch(probs, n, rules, nrules, table, nitems, maxit);
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
    if (mode == ST_BATCH_PERSISTENT) {
        int per_cu = 0, dev = 0, cus = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&k_batch_persistent), 256, 0);
        if (e != hipSuccess) return hip_check(e, "occupancy");
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int grid = per_cu * cus;
        const int need = (nitems + 3) / 4;
        if (grid > need) grid = need;
        if (grid < 1) grid = 1;
        int np = n, ni = nitems, ns = maxit;
        void *args[] = {&P, &np, &ni, &ns};
        if (elapsed_ms) (void)hipEventRecord(e0, s);
        e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_batch_persistent),
                                       dim3(grid), dim3(256), args, 0, s);
        if (e != hipSuccess) return hip_check(e, "batched cooperative launch");
    } else {
        // capture one launch per sweep, replay as one graph
        const dim3 grid((nitems + 3) / 4), block(256);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_check(e, "begin capture");
        for (int sw = 0; sw < maxit; ++sw)
            hipLaunchKernelGGL(k_batch_sweep, grid, block, 0, s, P, n, nitems, sw);
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
