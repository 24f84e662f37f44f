// COOP family: persistent cooperative kernel with grid-synchronised temporal
// blocking.
//
// Replaces coop (src/stencil_kernel_coop.cu:5-49), whose launch the
// reference comments out (src/stencil_dispatch.cu:160-177), and its
// grid-stride "all ITERATIONS in one kernel with grid.sync() per sweep"
// design.  Here one HBM pass advances the field by tb sweeps:
//
//   * a workgroup owns an output tile TX x TY (x, y) and a chunk of planes,
//     and marches along z;
//   * level 0 (the pass input) and every intermediate level t < tb keep a
//     ring of 2R+1 planes in LDS; level t covers the tile grown by R*(tb-t)
//     on each side in x and y (overlapped tiling, no exchange between
//     workgroups), and level tb is written to HBM;
//   * per input plane: one coalesced load of the level-0 region, then tb
//     LDS-to-LDS sweeps of one plane each, one barrier per level;
//   * passes are separated by a grid-wide barrier (cooperative launch) or
//     by kernel boundaries (one launch per pass).
//
// Boundary cells (the R-thick shell) of level t take their value from the
// buffer the per-sweep ping-pong would read at sweep s+t: u1 for even,
// u2 for odd global sweep index, exactly as the reference's alternating
// buffers provide it (both buffers' shells are never written).  The number
// of passes is chosen with the parity of the iteration count, so the result
// lands in the same buffer as with per-sweep launches.
//
// resolve_coop picks among the COOP forms: streaming rows (two fused sweeps
// per pass, sweep_stream.hip), register-queue frames (sweep_tblock.hip), the
// row kernel with a grid barrier after every sweep (sweep_rows.hip, the
// reference's own one-sweep-per-barrier schedule) and the LDS-ring tiles
// above.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "st_device.hpp"
#include "st_internal.hpp"

namespace st {

struct CoopArgs {
    Geom g;
    int TX, TY;          // output tile
    int zchunk;          // output planes per tile
    int tiles_x, tiles_y, ntiles;
    int iters;           // total sweeps
    int npass;           // passes
    int pass0, pass1;    // passes run by this launch
    int lds_t;           // level count the LDS layout was sized for
};

__device__ __forceinline__ void pass_range(int iters, int npass, int p, int &s, int &tb)
{
    const int q = iters / npass, rem = iters % npass;
    s = p * q + min(p, rem);
    tb = q + (p < rem ? 1 : 0);
}

template <int OP, int R, typename T>
__global__ __launch_bounds__(256) void k_coop(T *u1, T *u2, CoopArgs a, T kc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    constexpr int RING = 2 * R + 1;
    const Geom &g = a.g;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;

    for (int p = a.pass0; p < a.pass1; ++p) {
        int s, tb;
        pass_range(a.iters, a.npass, p, s, tb);
        const T *X = (p & 1) ? u2 : u1;  // pass input
        T *Y = (p & 1) ? u1 : u2;        // pass output
        // level frame: level-0 region of this pass
        const int W0 = a.TX + 2 * R * tb;
        const int H0 = a.TY + 2 * R * tb;
        const int P0 = W0 * H0;  // LDS plane size (elements)

        for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
            const int txi = tile % a.tiles_x;
            const int rest = tile / a.tiles_x;
            const int tyi = rest % a.tiles_y;
            const int tzi = rest / a.tiles_y;
            const int X0 = txi * a.TX, Y0 = tyi * a.TY;
            const int z0 = tzi * a.zchunk, z1 = min(z0 + a.zchunk, g.nz);
            const int fx = X0 - R * tb, fy = Y0 - R * tb;  // frame origin
            // Level t produces plane kin - t*R at step kin.  The march runs
            // R*tb steps past the last input plane so that every level also
            // reaches the top planes of its range; steps beyond nz-1 load
            // nothing at level 0.
            const int kin_lo = max(z0 - R * tb, 0), kin_hi = z1 + R * tb;

            for (int kin = kin_lo; kin < kin_hi; ++kin) {
                // ---- level 0: plane kin of the pass input (shell from B[s%2])
                if (kin < g.nz) {
                    const T *B = (s & 1) ? u2 : u1;
                    T *dst = lds + (kin % RING) * P0;
                    const bool kint = interior1(kin, g.nz, R);
                    const long long pz = (long long)kin * g.sz;
                    for (int e = tid; e < P0; e += nthr) {
                        const int ly = e / W0, lx = e - ly * W0;
                        const int x = fx + lx, y = fy + ly;
                        if (x < 0 || x >= g.nx || y < 0 || y >= g.ny) continue;
                        const long long idx = pz + (long long)y * g.sy + x;
                        const bool in_ = kint && interior1(x, g.nx, R) && interior1(y, g.ny, R);
                        // select the buffer, then load once (a select of two
                        // loads would fetch both)
                        const T *srcp = in_ ? X : B;
                        dst[e] = srcp[idx];
                    }
                }
                __syncthreads();
                // ---- levels 1..tb
                for (int t = 1; t <= tb; ++t) {
                    const int kt = kin - t * R;
                    const int grow = R * (tb - t);
                    const int lo = max(z0 - grow, 0), hi = min(z1 + grow, g.nz);
                    if (kt >= lo && kt < hi) {
                        const T *Bt = ((s + t) & 1) ? u2 : u1;
                        const T *src = lds + (t - 1) * RING * P0;
                        T *dst = lds + t * RING * P0 + (kt % RING) * P0;
                        const bool kint = interior1(kt, g.nz, R);
                        const long long pz = (long long)kt * g.sz;
                        // region of level t inside the frame
                        const int rx0 = R * t, ry0 = R * t;
                        const int rw = a.TX + 2 * grow, rh = a.TY + 2 * grow;
                        const int n = rw * rh;
                        const T *sc = src + (kt % RING) * P0;
                        for (int e = tid; e < n; e += nthr) {
                            const int ry = e / rw, rx = e - ry * rw;
                            const int lx = rx0 + rx, ly = ry0 + ry;
                            const int x = fx + lx, y = fy + ly;
                            if (x < 0 || x >= g.nx || y < 0 || y >= g.ny) continue;
                            const long long idx = pz + (long long)y * g.sy + x;
                            const int li = ly * W0 + lx;
                            T v;
                            if (kint && interior1(x, g.nx, R) && interior1(y, g.ny, R)) {
                                T u = 0;
#pragma unroll
                                for (int d = R; d >= 1; --d) u += sc[li - d];
#pragma unroll
                                for (int d = 1; d <= R; ++d) u += sc[li + d];
#pragma unroll
                                for (int d = R; d >= 1; --d) u += sc[li - d * W0];
#pragma unroll
                                for (int d = 1; d <= R; ++d) u += sc[li + d * W0];
#pragma unroll
                                for (int d = R; d >= 1; --d) u += src[((kt - d) % RING) * P0 + li];
#pragma unroll
                                for (int d = 1; d <= R; ++d) u += src[((kt + d) % RING) * P0 + li];
                                v = finish<OP>(u, sc[li], kc);
                                if (t == tb) {
                                    Y[idx] = v;
                                    continue;
                                }
                            } else {
                                if (t == tb) continue;  // shell is never written
                                v = Bt[idx];
                            }
                            dst[li] = v;
                        }
                    }
                    __syncthreads();
                }
            }
        }
        if (p + 1 < a.pass1) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- host side

template <int OP, int R, typename T>
static const void *coop_fn()
{
    return reinterpret_cast<const void *>(&k_coop<OP, R, T>);
}

static const void *pick_coop(const Cfg &c)
{
    const bool f64 = c.dtype == ST_F64;
#define CF(RR)                                                                    \
    case RR:                                                                      \
        if (f64) return c.op == OP_THESIS ? coop_fn<OP_THESIS, RR, double>()      \
                                          : coop_fn<OP_JACOBI, RR, double>();     \
        return c.op == OP_THESIS ? coop_fn<OP_THESIS, RR, float>()                \
                                 : coop_fn<OP_JACOBI, RR, float>();
    switch (c.r) {
        CF(1)
        CF(2)
        CF(3)
        CF(4)
    default: return nullptr;
    }
#undef CF
}

static int coop_npass(int iters, int tmax)
{
    if (iters <= 0) return 0;
    int n = (iters + tmax - 1) / tmax;
    if ((n & 1) != (iters & 1)) ++n;
    return n;
}

static long long coop_lds(const Cfg &c, int t)
{
    const int es = c.dtype == ST_F64 ? 8 : 4;
    const long long w0 = c.bx + 2LL * c.r * t, h0 = c.by + 2LL * c.r * t;
    return (long long)t * (2 * c.r + 1) * w0 * h0 * es;
}

// For COOP, block_x/block_y are the output tile (TX, TY), block_z the plane
// chunk, coarsen_y is unused and the workgroup is 256 threads.
// ---------------------------------------------------------------- grid-stride form
//
// The reference's coop kernel (src/stencil_kernel_coop.cu:5-49): persistent
// workgroups stride over the interior, a grid barrier between the sweeps.
// The unit of the stride is a group of four interior rows of one plane (one
// wave per row, its lanes 64 columns apart along the row, as the plain
// kernel's 64x4 blocks).  It is a coverage form, not a fast one: 29.5
// GCells/s at 512^3 R=8 against 107 for the plain kernel (31.6 with single
// cells as the unit, 23.6 with contiguous row ranges per workgroup).  The
// per-sweep ping-pong (sweep s reads buffer s%2) and the sum
// order of the other kernels (x, y, then z terms, nearest last on the minus
// side, nearest first on the plus side).  Any radius and either
// dimensionality: the COOP form for radius 5 and up, where the LDS-ring,
// register and row forms are not built, and for 2D grids the streaming form
// does not take.  Sweeps [s0, s1) of the run.
constexpr int kGsRows = 4;

template <int OP, int D, typename T>
__global__ __launch_bounds__(64 * kGsRows) void k_coop_gs(T *__restrict__ u1, T *__restrict__ u2, Geom g, int s0,
                                                          int s1, T kc)
{
    const int R = g.r;
    const int nxi = g.nx - 2 * R, nyi = g.ny - 2 * R, nzi = D == 3 ? g.nz - 2 * R : 1;
    const int ngy = nyi > 0 ? (nyi + kGsRows - 1) / kGsRows : 0;
    const long long ngroups = (nxi > 0 && nyi > 0 && nzi > 0) ? (long long)ngy * nzi : 0;
    const int lane = threadIdx.x & 63, wr = threadIdx.x >> 6;
    for (int s = s0; s < s1; ++s) {
        if (s > s0) cooperative_groups::this_grid().sync();
        const T *in = (s & 1) ? u2 : u1;
        T *out = (s & 1) ? u1 : u2;
        for (long long q = blockIdx.x; q < ngroups; q += gridDim.x) {
            const int gy = (int)(q % ngy);
            const int j = R + gy * kGsRows + wr;
            const int k = D == 3 ? R + (int)(q / ngy) : 0;
            if (j >= g.ny - R) continue;
            const long long row = (long long)j * g.sy + (long long)k * g.sz;
            for (int i = R + lane; i < g.nx - R; i += 64) {
                const long long idx = row + i;
                T u = 0;
                for (int d = R; d >= 1; --d) u += in[idx - d];
                for (int d = 1; d <= R; ++d) u += in[idx + d];
                for (int d = R; d >= 1; --d) u += in[idx - d * g.sy];
                for (int d = 1; d <= R; ++d) u += in[idx + d * g.sy];
                if constexpr (D == 3) {
                    for (int d = R; d >= 1; --d) u += in[idx - d * g.sz];
                    for (int d = 1; d <= R; ++d) u += in[idx + d * g.sz];
                }
                out[idx] = finish<OP>(u, in[idx], kc);
            }
        }
    }
}

template <typename T>
static const void *coop_gs_fn(int op, int dims)
{
    if (dims == 3)
        return op == OP_THESIS ? reinterpret_cast<const void *>(&k_coop_gs<OP_THESIS, 3, T>)
                               : reinterpret_cast<const void *>(&k_coop_gs<OP_JACOBI, 3, T>);
    return op == OP_THESIS ? reinterpret_cast<const void *>(&k_coop_gs<OP_THESIS, 2, T>)
                           : reinterpret_cast<const void *>(&k_coop_gs<OP_JACOBI, 2, T>);
}

static st_status resolve_coop_gs(Cfg &c)
{
    c.kind = 7;
    c.tsteps = 1;
    c.bx = 64;
    c.by = kGsRows;
    c.bz = 1;
    c.cx = c.cy = 1;
    c.zchunk = 0;
    c.lds = 0;
    c.gx = 0;  // decided at launch from occupancy
    c.gy = c.gz = 1;
    return ST_OK;
}

// Every sweep in one cooperative launch with a grid barrier between sweeps
// (the grid capped at the co-resident workgroups), or one launch per sweep
// with launch_per_pass or when the device refuses the cooperative launch.
static st_status run_coop_gs(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2)
{
    *in_u2 = c.iters & 1;
    if (c.iters <= 0) return ST_OK;
    const void *fn = c.dtype == ST_F64 ? coop_gs_fn<double>(c.op, c.dims) : coop_gs_fn<float>(c.op, c.dims);
    Geom g;
    g.nx = c.nx; g.ny = c.ny; g.nz = c.dims == 3 ? c.nz : 1; g.r = c.r;
    g.sy = c.nx; g.sz = (long long)c.nx * c.ny;
    const long long cells = (long long)std::max(c.nx - 2 * c.r, 0) * std::max(c.ny - 2 * c.r, 0) *
                            (c.dims == 3 ? std::max(c.nz - 2 * c.r, 0) : 1);
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * kGsRows, 0);
    if (e != hipSuccess) return hip_check(e, "occupancy query (grid-stride coop)");
    if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_check(e, "hipGetDevice");
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_check(e, "CU count");
    if (per_cu < 1 || cus < 1) return fail(ST_ELAUNCH, "grid-stride COOP kernel does not fit on a CU");
    // one workgroup per group of kGsRows interior rows, at most the co-resident ones
    long long grid = cells > 0 ? (long long)((c.ny - 2 * c.r + kGsRows - 1) / kGsRows) *
                                     (c.dims == 3 ? c.nz - 2 * c.r : 1)
                               : 0;
    if (grid > (long long)per_cu * cus) grid = (long long)per_cu * cus;
    if (grid < 1) grid = 1;
    double kd = c.op == OP_THESIS ? (double)(2 * c.dims * c.r) : c.weight;
    float kf = (float)kd;
    void *karg = c.dtype == ST_F64 ? static_cast<void *>(&kd) : static_cast<void *>(&kf);
    auto launch = [&](int s0, int s1, bool coop) {
        void *args[] = {&u1, &u2, &g, &s0, &s1, karg};
        return coop ? hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(64 * kGsRows), args, 0, s)
                    : hipLaunchKernel(fn, dim3((unsigned)grid), dim3(64 * kGsRows), args, 0, s);
    };
    if (!c.per_pass && c.iters > 1) {
        e = launch(0, c.iters, true);
        if (e == hipSuccess) return ST_OK;
        if (e != hipErrorIllegalState) return hip_check(e, "hipLaunchCooperativeKernel(grid-stride coop)");
        // the device takes no cooperative launch from this process now: the
        // same sweeps, one launch each
        (void)hipGetLastError();
    }
    for (int i = 0; i < c.iters; ++i)
        if ((e = launch(i, i + 1, false)) != hipSuccess) return hip_check(e, "grid-stride coop launch");
    return ST_OK;
}

st_status resolve_coop(Cfg &c)
{
    if (c.form == 5) return resolve_coop_gs(c);
    if (c.dims == 2) {
        // 2D: the streaming form (four or six fused sweeps per pass) where
        // it applies, else the grid-stride form
        if (c.form != 0 && c.form != 3)
            return fail(ST_EUNSUPPORTED, "COOP in 2D has the streaming and grid-stride forms (temporal_form 0, 3 "
                                         "or 5)");
        if (c.form == 0 && !stream2d_eligible(c)) return resolve_coop_gs(c);
        return resolve_stream2d(c);
    }
    if (c.dims != 3) return fail(ST_EUNSUPPORTED, "COOP is implemented for 2D and 3D grids");
    if (c.r > 4) {
        if (c.form != 0)
            return fail(ST_EUNSUPPORTED, "COOP forms 1-4 are built for radius 1..4; radius %d takes the "
                                         "grid-stride form (temporal_form 0 or 5)", c.r);
        return resolve_coop_gs(c);
    }
    const bool heur = c.heuristic || c.bx <= 0;
    // streaming form: requested, or the library's choice for heuristic
    // shapes where it applies (two fused sweeps)
    if (c.form == 3 || (c.form == 0 && heur && (c.tsteps == 0 || c.tsteps == 2) && stream_eligible(c))) {
        c.zchunk = (c.heuristic || c.bz <= 0) ? 0 : c.bz;
        if (c.heuristic) c.by = 0;
        return resolve_stream(c);
    }
    {
        // register-queue form: requested, or the library's choice for
        // heuristic shapes where it applies
        const int t = c.tsteps > 0 ? c.tsteps : 2;
        if (c.form == 2 || (c.form == 0 && heur && tblock_eligible(c, t))) {
            if (!tblock_eligible(c, t))
                return fail(ST_EUNSUPPORTED, "register temporal form needs FP32, radius 1, "
                                             "temporal_steps 2 and an even iteration count (T=%d)", t);
            if (heur) { c.cx = 0; c.cy = 0; c.zchunk = 0; }
            else c.zchunk = c.bz > 0 ? c.bz : 0;
            return resolve_tblock(c, t);
        }
    }
    // row form, one sweep per grid barrier: requested, or the library's
    // choice for heuristic shapes where it applies (every other radius and
    // FP64; measured on 512^3 FP32 R=2: 1.9 ms for 8 sweeps against 17 ms
    // for the LDS-ring tiles)
    if (c.form == 4 || (c.form == 0 && heur && c.tsteps <= 1 && rows_eligible(c))) return resolve_rows_coop(c);
    c.kind = 1;
    int want = c.tsteps > 0 ? c.tsteps : 2;
    if (want > c.iters && c.iters > 0) want = c.iters;
    if (want < 1) want = 1;
    if (heur) {
        // First tile (largest first) that holds `want` levels in LDS; else
        // the first that holds any.
        static const int tiles[][2] = {{64, 32}, {64, 16}, {32, 32}, {32, 16}, {16, 16}, {16, 8}, {8, 8}};
        int pick = -1, pick_t = 0;
        for (int i = 0; i < 7; ++i) {
            Cfg t = c;
            t.bx = tiles[i][0];
            t.by = tiles[i][1];
            int tt = want;
            while (tt > 1 && coop_lds(t, tt) > kMaxLds) --tt;
            if (coop_lds(t, tt) > kMaxLds) continue;
            if (tt == want) { pick = i; pick_t = tt; break; }
            if (pick < 0) { pick = i; pick_t = tt; }
        }
        if (pick < 0) return fail(ST_ELDS, "no COOP tile fits in LDS");
        c.bx = tiles[pick][0];
        c.by = tiles[pick][1];
        c.bz = 0;
        want = pick_t;
    }
    if (c.bx < 1 || c.by < 1) return fail(ST_EINVAL, "COOP tile must be positive");
    int t = want;
    while (t > 1 && coop_lds(c, t) > kMaxLds) --t;
    if (coop_lds(c, t) > kMaxLds)
        return fail(ST_ELDS, "COOP tile %dx%d needs %lld bytes of LDS", c.bx, c.by, coop_lds(c, t));
    c.tsteps = t;
    const int npass = coop_npass(c.iters, t);
    // the largest pass length actually used
    const int tmax_used = npass ? (c.iters + npass - 1) / npass : 1;
    c.lds = (int)coop_lds(c, tmax_used > 0 ? tmax_used : 1);
    const int tiles_x = (c.nx + c.bx - 1) / c.bx, tiles_y = (c.ny + c.by - 1) / c.by;
    int chunk = c.bz > 0 ? c.bz : 0;
    if (chunk <= 0) {
        const long long per = (long long)tiles_x * tiles_y;
        long long chunks = (1024 + per - 1) / per;
        chunk = (int)((c.nz + chunks - 1) / chunks);
        if (chunk < 32) chunk = 32;
        if (chunk > c.nz) chunk = c.nz;
    }
    c.zchunk = chunk;
    c.bz = chunk;
    c.cx = 1;
    c.cy = 1;
    c.gy = c.gz = 1;
    c.gx = 0;  // decided at launch from occupancy
    return ST_OK;
}

st_status run_coop(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2)
{
    if (c.kind == 2) return run_tblock(c, u1, u2, s, in_u2);
    if (c.kind == 3) return run_stream(c, u1, u2, s, in_u2);
    if (c.kind == 5) return run_rows_coop(c, u1, u2, s, in_u2);
    if (c.kind == 7) return run_coop_gs(c, u1, u2, s, in_u2);
    if (c.kind == 6) {
        // sweeps outside fused passes: the library's per-sweep 2D kernel
        Cfg sc = c;
        sc.variant = ST_V_SMEM_REGISTER;
        sc.heuristic = 1;
        sc.bx = sc.by = sc.bz = 0;
        sc.cx = 1;
        sc.cy = 0;
        sc.form = 0;
        sc.tsteps = 0;
        sc.zchunk = 0;
        st_status st = resolve_march(sc);
        if (st != ST_OK) return st;
        return run_stream2d(c, sc, u1, u2, s, in_u2);
    }
    const void *fn = pick_coop(c);
    if (!fn) return fail(ST_EUNSUPPORTED, "no COOP kernel for radius %d", c.r);
    CoopArgs a;
    a.g.nx = c.nx; a.g.ny = c.ny; a.g.nz = c.nz; a.g.r = c.r;
    a.g.sy = c.nx; a.g.sz = (long long)c.nx * c.ny;
    a.TX = c.bx; a.TY = c.by; a.zchunk = c.zchunk;
    a.tiles_x = (c.nx + c.bx - 1) / c.bx;
    a.tiles_y = (c.ny + c.by - 1) / c.by;
    a.ntiles = a.tiles_x * a.tiles_y * ((c.nz + c.zchunk - 1) / c.zchunk);
    a.iters = c.iters;
    a.npass = coop_npass(c.iters, c.tsteps);
    a.lds_t = c.tsteps;
    *in_u2 = a.npass & 1;
    if (a.npass == 0) return ST_OK;
    const int threads = 256;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
    if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(COOP LDS)");
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, c.lds);
    if (e != hipSuccess) return hip_check(e, "occupancy query");
    if (per_cu < 1) return fail(ST_ELDS, "COOP kernel does not fit on a CU (lds %d)", c.lds);
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int grid = per_cu * cus;
    if (grid > a.ntiles) grid = a.ntiles;
    const double kcd = c.op == OP_THESIS ? (double)(2 * c.dims * c.r) : c.weight;
    float kf = (float)kcd;
    double kd = kcd;
    void *karg = c.dtype == ST_F64 ? (void *)&kd : (void *)&kf;
    bool coop = !c.per_pass;
    if (coop) {
        a.pass0 = 0;
        a.pass1 = a.npass;
        void *args[] = {&u1, &u2, &a, karg};
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, c.lds, s);
        if (e == hipErrorIllegalState) {
            // the device takes no cooperative launch from this process now:
            // the same passes, one launch each
            (void)hipGetLastError();
            coop = false;
        } else if (e != hipSuccess) {
            return hip_check(e, "hipLaunchCooperativeKernel");
        }
    }
    if (!coop) {
        for (int p = 0; p < a.npass; ++p) {
            a.pass0 = p;
            a.pass1 = p + 1;
            void *args[] = {&u1, &u2, &a, karg};
            e = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, c.lds, s);
            if (e != hipSuccess) return hip_check(e, "COOP pass launch");
        }
    }
    return ST_OK;
}

}  // namespace st
