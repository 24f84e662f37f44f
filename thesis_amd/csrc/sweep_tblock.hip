// COOP family, register-queue form: temporal blocking with full-row frames.
//
// Same contract as k_coop in sweep_coop.hip (passes of tb fused sweeps,
// boundary shell of level t read from the buffer the per-sweep ping-pong
// would hold at sweep s+t, pass count with the parity of the iteration
// count), organised for HBM streaming:
//
//   * a workgroup of WX*WY waves owns a frame of WX*64*CX columns by
//     WY*RPW rows and marches it along z; when the frame is at least as wide
//     as the grid it spans whole rows and needs no x halo, otherwise frames
//     overlap by R*T columns on each side;
//   * every thread owns CX consecutive cells of RPW consecutive rows and
//     keeps, for each time level t < T, a register queue of the last 2R+1
//     planes of those cells (z neighbours never touch LDS);
//   * per input plane: the next plane's cells are prefetched into registers
//     before the current plane is processed; for each level the centre plane
//     of the level below is written to one LDS plane, one barrier, and x/y
//     neighbours are read back from LDS (own registers where the neighbour
//     is in the thread's own cells);
//   * only level T is stored, for the frame core (frame shrunk by R*T in y,
//     and in x when frames overlap).
// The LDS plane of a level is rewritten one step later, with at least one
// barrier of the other levels in between, so T >= 2 needs no second buffer.
#include <hip/hip_cooperative_groups.h>

#include "st_device.hpp"
#include "st_internal.hpp"

namespace st {

struct TbArgs {
    Geom g;
    int wx, wy;            // waves along x and y
    int fw, fwp;           // frame width, LDS row pitch
    int fh;                // frame rows (wy * RPW)
    int core_w, core_h;    // output core of a frame
    int xfull;             // frame spans whole rows (no x halo)
    int zchunk;
    int tiles_x, tiles_y, ntiles;
    int iters, npass, pass0, pass1;
};

__device__ __forceinline__ void tb_pass_range(int iters, int npass, int p, int &s, int &tb)
{
    const int q = iters / npass, rem = iters % npass;
    s = p * q + min(p, rem);
    tb = q + (p < rem ? 1 : 0);
}

template <int OP, int R, int T, int CX, int RPW>
__global__ __launch_bounds__(1024) void k_tblock(float *u1, float *u2, TbArgs a, float kc)
{
    static_assert(T >= 2, "one LDS plane per level needs T >= 2");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float *lds = reinterpret_cast<float *>(lds_raw);
    constexpr int RING = 2 * R + 1;
    const Geom &g = a.g;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wxi = wave % a.wx, wyi = wave / a.wx;
    const int xl0 = (wxi * 64 + lane) * CX;  // frame-local x of cell 0
    const int yl0 = wyi * RPW;               // frame-local row of row 0
    const int plane = a.fh * a.fwp;
    const unsigned nblk = gridDim.x;
    const unsigned bid = xcd_remap(blockIdx.x, nblk);

    for (int p = a.pass0; p < a.pass1; ++p) {
        int s, tb;
        tb_pass_range(a.iters, a.npass, p, s, tb);
        // this kernel is instantiated for exactly T fused sweeps; shorter
        // passes are not routed here
        const float *X = (p & 1) ? u2 : u1;
        float *Y = (p & 1) ? u1 : u2;
        const float *B0 = (s & 1) ? u2 : u1;

        for (int tile = (int)bid; tile < a.ntiles; tile += (int)nblk) {
            const int txi = tile % a.tiles_x;
            const int rest = tile / a.tiles_x;
            const int tyi = rest % a.tiles_y;
            const int tzi = rest / a.tiles_y;
            const int fx0 = a.xfull ? 0 : txi * a.core_w - R * T;
            const int fy0 = tyi * a.core_h - R * T;
            const int z0 = tzi * a.zchunk, z1 = min(z0 + a.zchunk, g.nz);
            // output core in frame-local coordinates
            const int cx0 = a.xfull ? 0 : R * T, cx1 = a.xfull ? a.fw : R * T + a.core_w;
            const int cy0 = R * T, cy1 = R * T + a.core_h;

            // per-cell coordinates and flags
            int gx[CX], gy[RPW];
            bool xin[CX], xint[CX], yin[RPW], yint[RPW], xcore[CX], ycore[RPW];
#pragma unroll
            for (int c = 0; c < CX; ++c) {
                gx[c] = fx0 + xl0 + c;
                xin[c] = gx[c] >= 0 && gx[c] < g.nx && xl0 + c < a.fw;
                xint[c] = interior1(gx[c], g.nx, R);
                xcore[c] = xl0 + c >= cx0 && xl0 + c < cx1;
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                gy[r] = fy0 + yl0 + r;
                yin[r] = gy[r] >= 0 && gy[r] < g.ny;
                yint[r] = interior1(gy[r], g.ny, R);
                ycore[r] = yl0 + r >= cy0 && yl0 + r < cy1;
            }
            bool all_xin = true;
#pragma unroll
            for (int c = 0; c < CX; ++c) all_xin = all_xin && xin[c];
            const int gxl = min(max(gx[0], 0), g.nx - CX);  // clamped vector column
            const bool xvec = (gx[0] & (CX - 1)) == 0;      // 4*CX-byte aligned cells

            float q[T][RING][RPW][CX];  // level queues: q[t][slot][r][c]
            float pf[RPW][CX];          // prefetched level-0 plane

            auto load0 = [&](int kp, float (&dst)[RPW][CX]) {
                // level-0 plane kp: interior from X, shell from B0
                if (kp < 0 || kp >= g.nz) return;
                const bool kint = interior1(kp, g.nz, R);
                const long long pz = (long long)kp * g.sz;
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    if (!yin[r]) continue;
                    const long long rb = pz + (long long)gy[r] * g.sy;
                    const bool rint = kint && yint[r];
                    if (xvec && all_xin && rint && xint[0] && xint[CX - 1]) {
                        typedef float V __attribute__((ext_vector_type(CX)));
                        const V v = *reinterpret_cast<const V *>(X + rb + gxl);
#pragma unroll
                        for (int c = 0; c < CX; ++c) dst[r][c] = v[c];
                    } else {
#pragma unroll
                        for (int c = 0; c < CX; ++c) {
                            if (!xin[c]) continue;
                            const float *src = (rint && xint[c]) ? X : B0;
                            dst[r][c] = src[rb + gx[c]];
                        }
                    }
                }
            };

            const int kin_lo = z0 - R * T, kin_hi = z1 + R * T;
            load0(kin_lo, pf);
            for (int kb = kin_lo; kb < kin_hi; kb += RING) {
#pragma unroll
                for (int ph = 0; ph < RING; ++ph) {
                    const int kin = kb + ph;
                    if (kin >= kin_hi) break;
                    // level 0 newest plane = kin (slot ph), prefetch kin+1
#pragma unroll
                    for (int r = 0; r < RPW; ++r)
#pragma unroll
                        for (int c = 0; c < CX; ++c) q[0][ph][r][c] = pf[r][c];
                    load0(kin + 1, pf);
                    // levels 1..T
#pragma unroll
                    for (int t = 1; t <= T; ++t) {
                        const int kt = kin - t * R;         // plane produced now
                        // centre plane of level t-1 sits R planes behind its newest
                        const int cs = (ph - R + RING) % RING;  // slot of plane kin-(t-1)R-R
                        float(&cen)[RPW][CX] = q[t - 1][cs];
                        float *L = lds + (t - 1) * plane;
                        // publish the centre plane
#pragma unroll
                        for (int r = 0; r < RPW; ++r) {
                            typedef float V __attribute__((ext_vector_type(CX)));
                            V v;
#pragma unroll
                            for (int c = 0; c < CX; ++c) v[c] = cen[r][c];
                            *reinterpret_cast<V *>(L + (yl0 + r) * a.fwp + xl0) = v;
                        }
                        __syncthreads();
                        const bool kin_dom = kt >= 0 && kt < g.nz;
                        const bool kint = interior1(kt, g.nz, R);
                        const float *Bt = ((s + t) & 1) ? u2 : u1;
                        const long long pz = (long long)kt * g.sz;
                        float res[RPW][CX];
#pragma unroll
                        for (int r = 0; r < RPW; ++r) {
                            const int ly = yl0 + r;
#pragma unroll
                            for (int c = 0; c < CX; ++c) {
                                const int lx = xl0 + c;
                                float u = 0.0f;
#pragma unroll
                                for (int d = R; d >= 1; --d)
                                    u += (c - d >= 0) ? cen[r][c - d] : L[ly * a.fwp + max(lx - d, 0)];
#pragma unroll
                                for (int d = 1; d <= R; ++d)
                                    u += (c + d < CX) ? cen[r][c + d]
                                                      : L[ly * a.fwp + min(lx + d, a.fwp - 1)];
#pragma unroll
                                for (int d = R; d >= 1; --d)
                                    u += (r - d >= 0) ? cen[r - d][c] : L[max(ly - d, 0) * a.fwp + lx];
#pragma unroll
                                for (int d = 1; d <= R; ++d)
                                    u += (r + d < RPW) ? cen[r + d][c]
                                                       : L[min(ly + d, a.fh - 1) * a.fwp + lx];
#pragma unroll
                                for (int d = R; d >= 1; --d) u += q[t - 1][(cs - d + RING) % RING][r][c];
#pragma unroll
                                for (int d = 1; d <= R; ++d) u += q[t - 1][(cs + d) % RING][r][c];
                                res[r][c] = finish<OP>(u, cen[r][c], kc);
                            }
                        }
                        // shell cells take the boundary value of level t
                        if (kin_dom) {
#pragma unroll
                            for (int r = 0; r < RPW; ++r) {
                                if (!yin[r]) continue;
                                const bool rint = kint && yint[r];
                                if (rint && xint[0] && xint[CX - 1]) continue;
                                const long long rb = pz + (long long)gy[r] * g.sy;
#pragma unroll
                                for (int c = 0; c < CX; ++c)
                                    if (xin[c] && !(rint && xint[c])) res[r][c] = Bt[rb + gx[c]];
                            }
                        }
                        if (t < T) {
                            // newest plane of level t goes to slot ph of its queue
#pragma unroll
                            for (int r = 0; r < RPW; ++r)
#pragma unroll
                                for (int c = 0; c < CX; ++c) q[t][ph][r][c] = res[r][c];
                        } else if (kt >= z0 && kt < z1 && kint) {
#pragma unroll
                            for (int r = 0; r < RPW; ++r) {
                                if (!(ycore[r] && yin[r] && yint[r])) continue;
                                float *row = Y + pz + (long long)gy[r] * g.sy;
                                bool allc = all_xin && xvec;
#pragma unroll
                                for (int c = 0; c < CX; ++c) allc = allc && xcore[c] && xint[c];
                                if (allc) {
                                    typedef float V __attribute__((ext_vector_type(CX)));
                                    V v;
#pragma unroll
                                    for (int c = 0; c < CX; ++c) v[c] = res[r][c];
                                    __builtin_nontemporal_store(v, reinterpret_cast<V *>(row + gx[0]));
                                } else {
#pragma unroll
                                    for (int c = 0; c < CX; ++c)
                                        if (xin[c] && xcore[c] && xint[c]) row[gx[c]] = res[r][c];
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();  // LDS reuse by the next tile
        }
        if (p + 1 < a.pass1) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- host side

struct TbEntry {
    int r, t, cx, rpw;
    const void *fn[2];
};

template <int OP, int R, int T, int CX, int RPW>
static const void *tb_fn()
{
    return reinterpret_cast<const void *>(&k_tblock<OP, R, T, CX, RPW>);
}

#define TBE(R, T, CX, RPW) \
    {R, T, CX, RPW, {tb_fn<OP_THESIS, R, T, CX, RPW>(), tb_fn<OP_JACOBI, R, T, CX, RPW>()}}
// Instantiations that fit 128 VGPRs (16 waves per CU) without spilling.
static const TbEntry kTb[] = {
    TBE(1, 2, 2, 4), TBE(1, 2, 4, 2),
};
#undef TBE

static const TbEntry *find_tb(int r, int t, int cx, int rpw)
{
    for (const TbEntry &e : kTb)
        if (e.r == r && e.t == t && e.cx == cx && e.rpw == rpw) return &e;
    return nullptr;
}

// Register-queue form applies to FP32, radius 1, two fused sweeps, and iteration counts that
// split into passes of exactly T sweeps with the right parity.
bool tblock_eligible(const Cfg &c, int t)
{
    if (c.dtype != ST_F32 || c.dims != 3 || t < 2 || c.iters <= 0 || c.nx % 2) return false;
    if (c.iters % t) return false;
    const int npass = c.iters / t;
    if ((npass & 1) != (c.iters & 1)) return false;
    for (const TbEntry &e : kTb)
        if (e.r == c.r && e.t == t) return true;
    return false;
}

st_status resolve_tblock(Cfg &c, int t)
{
    // defaults: 2 values per lane, 6 rows per wave, 4 waves across 512 columns
    int cx = c.cx > 1 ? c.cx : 2, rpw = c.cy > 0 ? c.cy : 4;
    while (cx > 2 && c.nx % cx) cx /= 2;  // vector cells must not straddle rows
    const TbEntry *e = find_tb(c.r, t, cx, rpw);
    if (!e) e = find_tb(c.r, t, 2, 4);
    if (!e) for (const TbEntry &x : kTb) if (x.r == c.r && x.t == t) { e = &x; break; }
    if (!e) return fail(ST_EUNSUPPORTED, "no register temporal kernel for r=%d T=%d", c.r, t);
    cx = e->cx;
    rpw = e->rpw;
    if (c.nx % cx) return fail(ST_EUNSUPPORTED, "nx must be a multiple of %d", cx);
    // waves across x: whole rows when 64*cx*wx >= nx within 1024 threads
    int wx = (c.nx + 64 * cx - 1) / (64 * cx);
    int wy;
    bool xfull = true;
    if (wx > 8) { wx = 4; xfull = false; }
    wy = 16 / wx;
    if (wy > 4) wy = 4;
    if (wy < 1) wy = 1;
    const int fw = wx * 64 * cx;
    const int fh = wy * rpw;
    const int core_h = fh - 2 * c.r * t;
    const int core_w = xfull ? fw : fw - 2 * c.r * t;
    if (core_h < 1 || core_w < 1) return fail(ST_EUNSUPPORTED, "frame too small for T=%d", t);
    c.wx = wx;
    c.wy = wy;
    c.kind = 2;
    c.cx = cx;
    c.cy = rpw;
    c.bx = xfull ? c.nx : core_w;
    c.by = core_h;
    c.tsteps = t;
    const int fwp = fw;  // rows of 4*fw bytes: consecutive rows hit different banks
    c.lds = t * fh * fwp * 4;
    if (c.lds > kMaxLds) return fail(ST_ELDS, "register temporal frame needs %d bytes of LDS", c.lds);
    const int tiles_x = xfull ? 1 : (c.nx + core_w - 1) / core_w;
    const int tiles_y = (c.ny + core_h - 1) / core_h;
    int chunk = c.zchunk;
    if (chunk <= 0) {
        // about two frames per CU, chunks no thinner than 8*R*T planes
        const long long per = (long long)tiles_x * tiles_y;
        long long chunks = (512 + per - 1) / per;
        chunk = (int)((c.nz + chunks - 1) / chunks);
        if (chunk < 8 * c.r * t) chunk = 8 * c.r * t;
        if (chunk > c.nz) chunk = c.nz;
    }
    c.zchunk = chunk;
    c.bz = chunk;
    c.gx = 0;
    c.gy = c.gz = 1;
    return ST_OK;
}

st_status run_tblock(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2)
{
    const TbEntry *e = find_tb(c.r, c.tsteps, c.cx, c.cy);
    if (!e) return fail(ST_EUNSUPPORTED, "register temporal kernel missing");
    const void *fn = e->fn[c.op];
    TbArgs a;
    a.g.nx = c.nx; a.g.ny = c.ny; a.g.nz = c.nz; a.g.r = c.r;
    a.g.sy = c.nx; a.g.sz = (long long)c.nx * c.ny;
    a.wx = c.wx; a.wy = c.wy;
    a.fw = c.wx * 64 * c.cx;
    a.fwp = a.fw;
    a.fh = c.wy * c.cy;
    a.xfull = c.bx == c.nx && a.fw >= c.nx;
    a.core_w = a.xfull ? a.fw : a.fw - 2 * c.r * c.tsteps;
    a.core_h = a.fh - 2 * c.r * c.tsteps;
    a.zchunk = c.zchunk;
    a.tiles_x = a.xfull ? 1 : (c.nx + a.core_w - 1) / a.core_w;
    a.tiles_y = (c.ny + a.core_h - 1) / a.core_h;
    a.ntiles = a.tiles_x * a.tiles_y * ((c.nz + c.zchunk - 1) / c.zchunk);
    a.iters = c.iters;
    a.npass = c.iters / c.tsteps;
    *in_u2 = a.npass & 1;
    const int threads = 64 * c.wx * c.wy;
    hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
    if (err != hipSuccess) return hip_check(err, "hipFuncSetAttribute(tblock LDS)");
    int per_cu = 0;
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, c.lds);
    if (err != hipSuccess) return hip_check(err, "occupancy query");
    if (per_cu < 1) return fail(ST_ELDS, "register temporal kernel does not fit on a CU");
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int grid = per_cu * cus;
    if (grid > a.ntiles) grid = a.ntiles;
    float kf = c.op == OP_THESIS ? (float)(2 * c.dims * c.r) : (float)c.weight;
    bool coop = !c.per_pass;
    if (coop) {
        a.pass0 = 0;
        a.pass1 = a.npass;
        void *args[] = {&u1, &u2, &a, &kf};
        err = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, c.lds, s);
        if (err == hipErrorIllegalState) {
            // no cooperative launch from this process now: one launch per pass
            (void)hipGetLastError();
            coop = false;
        } else if (err != hipSuccess) {
            return hip_check(err, "hipLaunchCooperativeKernel(tblock)");
        }
    }
    if (!coop) {
        for (int p = 0; p < a.npass; ++p) {
            a.pass0 = p;
            a.pass1 = p + 1;
            void *args[] = {&u1, &u2, &a, &kf};
            err = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, c.lds, s);
            if (err != hipSuccess) return hip_check(err, "tblock pass launch");
        }
    }
    return ST_OK;
}

}  // namespace st
