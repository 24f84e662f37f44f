// COOP family, streaming form: two fused sweeps per HBM pass, whole grid
// rows per wave.
//
// Contract (as the other COOP forms): `iters` sweeps of the per-sweep
// ping-pong (sweep s reads buffer s%2, writes the interior of the other),
// result in u2 exactly when iters is odd.  Both buffers' outermost shells
// are compared on the GPU first (k_shell_eq):
//
//   * equal shells (ST_INIT_SEEDED, or any state where both buffers hold the
//     same boundary): a shell cell has the same value at every time level,
//     so two sweeps fuse into one pass with no boundary reloads.  Passes of
//     two sweeps run first, then (iters mod 4) single-sweep passes, which
//     keeps the pass count's parity equal to the iteration count's;
//   * different shells (ST_INIT_REFERENCE leaves u2 zero): the passes fuse
//     two sweeps as well.  Level 0 of a fused pass is the state after an
//     even number of sweeps, whose shell is u1's; level 1 has u2's shell.
//     The pass reads the interior from the buffer X it streams, but takes
//     every level-0 shell value from u1 and every level-1 shell value from
//     u2: z-shell planes and y-shell rows come from those buffers directly,
//     the x-shell cells (columns 0 and nx-1 of each row) from a compact
//     table of both buffers' x-shell values that k_shell_eq writes, so the
//     pass moves the same HBM bytes as with equal shells.  Only interior
//     cells are stored, so both shells stay untouched.
//
// Layout: a workgroup of WY wave rows owns a frame of RPW*WY full rows (a
// wave row is WX waves side by side; each wave holds RPW rows, each lane CX
// cells of each, in quads of four consecutive columns 256 apart: two rows
// of 4 or 8 cells up to 512 columns; over 512 columns two waves of 8 cells
// per row, or one wave of 16 cells on request) and marches it along z over
// a chunk of planes.  Per level and plane:
//   x neighbours   own registers, lane edges by DPP wave shifts; with two
//                  waves per row the seam columns (one value per row and
//                  wave) go through LDS with the y exchange;
//   y neighbours   own registers, frame-row edges through one LDS exchange
//                  of both rows (one barrier);
//   z neighbours   a running partial sum: when plane k arrives its x, y
//                  and k-1 terms are added (in the reference order) and the
//                  k+1 term completes it one plane later.
// Level 0 of the next plane is prefetched into registers before the current
// plane is processed; global loads stay in flight across the LDS barriers.
// Frames overlap by two rows on each side (one per fused sweep) and only
// their core rows are stored; z chunks overlap by two planes likewise.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "st_device.hpp"
#include "st_internal.hpp"

namespace st {

constexpr int kShellBlocks = 256;

struct SmArgs {
    int nx, ny, nz;
    long long sz;         // nx*ny
    int core_h;           // stored rows per frame (frame rows - 4)
    int tiles_y, zchunk, ntiles;
    int iters;
    int n2, n1;           // passes of two sweeps, then of one
    int p0, p1;           // passes of this launch: [p0, min(p1, passes))
    const int *shell_eq;  // kShellBlocks flags from k_shell_eq
    const float *xs;      // x-shell table of both buffers (xshell_index)
};

// Index of the x-shell value of buffer b (0: u1, 1: u2) at row j of plane
// k, column 0 (e = 0) or nx-1 (e = 1), in the table k_shell_eq writes.
// Planes are contiguous per row: a wave marching z reads one 128-byte line
// per 16 planes of each of its rows (plane-major order measured 20% slower
// with different shells: a new line, mostly an HBM miss, every plane).
__device__ __forceinline__ long long xshell_index(int b, int k, int j, int e, int ny, int nz)
{
    return (((long long)b * ny + j) * nz + k) * 2 + e;
}

// part[b] = 1 when every shell cell block b visits holds the same bits in
// u1 and u2; the x-shell cells of both buffers (columns 0 and nx-1 of every
// row) are copied into the table xs on the way.
__global__ __launch_bounds__(256) void k_shell_eq(const float *__restrict__ u1,
                                                  const float *__restrict__ u2, int nx, int ny,
                                                  int nz, int *__restrict__ part, float *__restrict__ xs)
{
    const long long A = (long long)nx * ny;
    const long long P = 2LL * nx + 2LL * (ny - 2);  // shell cells of an inner plane
    const long long n = 2 * A + (long long)(nz - 2) * P;
    int ok = 1;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
        long long idx;
        int x, y, z;
        if (i < 2 * A) {
            const long long j = i < A ? i : i - A;
            z = i < A ? 0 : nz - 1;
            y = (int)(j / nx);
            x = (int)(j - (long long)y * nx);
            idx = (long long)z * A + j;
        } else {
            const long long q = i - 2 * A;
            z = 1 + (int)(q / P);
            long long j = q % P;
            if (j < 2LL * nx) {
                y = j < nx ? 0 : ny - 1;
                x = (int)(j < nx ? j : j - nx);
            } else {
                j -= 2LL * nx;
                y = 1 + (int)(j >> 1);
                x = (j & 1) ? nx - 1 : 0;
            }
            idx = (long long)z * A + (long long)y * nx + x;
        }
        const float a = u1[idx], b = u2[idx];
        ok &= __float_as_uint(a) == __float_as_uint(b);
        if (x == 0 || x == nx - 1) {
            const int e = x == 0 ? 0 : 1;
            xs[xshell_index(0, z, y, e, ny, nz)] = a;
            xs[xshell_index(1, z, y, e, ny, nz)] = b;
        }
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) part[blockIdx.x] = ok;
}

// Lane i receives v of lane i-1 / i+1, rotating at the wave ends (wave-wide
// DPP rotates; every lane reads a valid lane, so bound_ctrl never applies and
// only spares the copy of the old value).
__device__ __forceinline__ float from_prev_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x13C, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_next_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, true));
}

// A lane's cells of one row: CX = 4*H values, quad h holding columns
// 256*h + 4*lane + 0..3, so that one 16-byte access per lane covers 1 KiB of
// the row contiguously (whole cache lines on every load and store).
template <int CX, int RPW>
struct SmPlane {
    float v[RPW][CX];
};

// Per-lane column constants.
template <int CX>
struct SmLane {
    int xo[CX / 4];    // first column of quad h, clamped into the row
    bool act[CX / 4];  // quad h lies inside the row
    bool lo[CX / 4];   // quad h starts at column 0
    bool hi[CX / 4];   // quad h ends at column nx-1
    bool whole[CX / 4];  // quad h is stored as one vector (interior, or any
                         // quad inside the row when both shells are equal)
};

// Per-tile, per-wave constants.
template <int RPW>
struct SmTile {
    int rowoff[RPW];            // clamped row offsets within a plane
    int rowc[RPW];              // clamped row index
    bool yshell[RPW];           // row outside the interior (wave-uniform)
    bool ytrue[RPW];            // row 0 or ny-1 of the grid (wave-uniform)
    bool ystore[RPW];           // row is an interior core row (wave-uniform)
    int z0, z1;
};

// Shell values of a fused pass with different shells (unused with equal
// shells).  The x-shell values (columns 0 and nx-1) of the tile's frame
// rows, for every plane the tile touches and both buffers, are copied from
// the table into LDS when the tile starts (SmXSlice); the pass reads them
// there (level 0: u1's, patched in when a plane is consumed; level 1:
// u2's).  Loading them from the table every plane instead (four small
// loads per plane, with their address arithmetic) made the pass 20%
// slower.  SmShell1 holds the grid's y-shell rows of u2 for the plane that
// the next step finishes at level 1 (rows with ytrue only).
template <int CX, int RPW>
struct SmShell1 {
    float y1[RPW][CX];  // u2, rows with ytrue
};

// LDS copy of the tile's x-shell values: entry (b, fr, kk) holds buffer b's
// values at frame row fr (clamped into the grid) of plane z0-3+kk (clamped).
struct SmXSlice {
    float2 *v;
    int rows;    // frame rows
    int stride;  // entries per row: zchunk + 5
    int k0;      // plane of entry 0: z0 - 3
    __device__ float2 at(int b, int fr, int k) const { return v[(b * rows + fr) * stride + (k - k0)]; }
};

// Level-0 plane k (clamped into the grid) from X.  DIFF: shell planes and
// shell rows from S0 (u1) instead; the x-shell values of u1 are patched in
// when the plane is consumed (sm_take).
template <int CX, int RPW, bool DIFF>
__device__ __forceinline__ void sm_load(SmPlane<CX, RPW> &d, const float *__restrict__ X,
                                        const float *__restrict__ S0, const SmArgs &a, const SmTile<RPW> &t,
                                        const SmLane<CX> &l, int k)
{
    typedef float V4 __attribute__((ext_vector_type(4)));
    k = min(max(k, 0), a.nz - 1);
    const long long po = (long long)k * a.sz;
    const float *base = (DIFF && (k == 0 || k == a.nz - 1) ? S0 : X) + po;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const float *rb = (DIFF && t.yshell[r] ? S0 + po : base) + t.rowoff[r];
#pragma unroll
        for (int h = 0; h < CX / 4; ++h) {
            const V4 q = *reinterpret_cast<const V4 *>(rb + l.xo[h]);
            d.v[r][4 * h + 0] = q[0];
            d.v[r][4 * h + 1] = q[1];
            d.v[r][4 * h + 2] = q[2];
            d.v[r][4 * h + 3] = q[3];
        }
    }
}

// Level-1 y-shell rows of plane k from u2 (S1).
template <int CX, int RPW>
__device__ __forceinline__ void sm_load_sh1(SmShell1<CX, RPW> &sh, const float *__restrict__ S1,
                                            const SmArgs &a, const SmTile<RPW> &t, const SmLane<CX> &l, int k)
{
    typedef float V4 __attribute__((ext_vector_type(4)));
    k = min(max(k, 0), a.nz - 1);
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        if (t.ytrue[r]) {
            const float *rb = S1 + (long long)k * a.sz + t.rowoff[r];
#pragma unroll
            for (int h = 0; h < CX / 4; ++h) {
                const V4 q = *reinterpret_cast<const V4 *>(rb + l.xo[h]);
#pragma unroll
                for (int j = 0; j < 4; ++j) sh.y1[r][4 * h + j] = q[j];
            }
        }
    }
}

// The loaded plane k with its level-0 x-shell cells replaced by u1's
// (DIFF; frame rows w*RPW + r).
template <int CX, int RPW, bool DIFF>
__device__ __forceinline__ void sm_take(SmPlane<CX, RPW> &d, const SmPlane<CX, RPW> &p, const SmXSlice &xsl,
                                        int w, int k, const SmLane<CX> &l)
{
    d = p;
    if constexpr (DIFF) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const float2 x0 = xsl.at(0, w * RPW + r, k);
#pragma unroll
            for (int h = 0; h < CX / 4; ++h) {
                d.v[r][4 * h] = l.lo[h] ? x0.x : d.v[r][4 * h];
                d.v[r][4 * h + 3] = l.hi[h] ? x0.y : d.v[r][4 * h + 3];
            }
        }
    }
}

// Partial sums of the plane `nw` (x, y, then the z-previous plane `zm`) in
// the reference's order; the y edges come from the neighbouring wave rows
// through `reg` (one LDS region of 2*waves*CX/4*64 float4).  With two waves
// per row (WX = 2, RPW = 2: both rows of a wave are edge rows, so the
// region holds them whole) the seam column next to a wave's first or last
// column is read from the partner wave's rows there.  w: wave index in the
// workgroup, wy: wave rows.
template <int CX, int RPW, int WX>
__device__ __forceinline__ void sm_partial(SmPlane<CX, RPW> &out, const SmPlane<CX, RPW> &nw,
                                           const SmPlane<CX, RPW> &zm, float4 *reg, int w, int wy, int lane)
{
    static_assert(WX == 1 || (WX == 2 && RPW == 2), "one wave per row, or two with two rows each");
    constexpr int H = CX / 4;
    const int wr = w / WX, wx = w % WX;
    // edge rows only: row 0 (e = 0) and row RPW-1 (e = 1)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int r = e ? RPW - 1 : 0;
#pragma unroll
        for (int h = 0; h < H; ++h)
            reg[((w * 2 + e) * H + h) * 64 + lane] =
                make_float4(nw.v[r][4 * h], nw.v[r][4 * h + 1], nw.v[r][4 * h + 2], nw.v[r][4 * h + 3]);
    }
    __syncthreads();
    // wrapped rows are outside the frame core and never stored
    const int wm = (wr == 0 ? wy - 1 : wr - 1) * WX + wx;
    const int wp = (wr == wy - 1 ? 0 : wr + 1) * WX + wx;
    // the partner wave's seam cell of row r (edge row e = r): the left
    // half's last column (quad H-1, lane 63, .w) or the right half's first
    // (quad 0, lane 0, .x)
    // (one value per row for the whole wave, read first and held in scalar
    // registers, so that it holds no vector register while the y edges are
    // read)
    float sc[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) sc[r] = 0.0f;
    if constexpr (WX == 2) {
        const float *seam = reinterpret_cast<const float *>(reg) +
                            4 * (wx == 1 ? ((w - 1) * 2 * H + H - 1) * 64 + 63 : (w + 1) * 2 * H * 64) +
                            (wx == 1 ? 3 : 0);
#pragma unroll
        for (int r = 0; r < RPW; ++r)
            sc[r] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(seam[r * H * 64 * 4])));
        __builtin_amdgcn_sched_barrier(0);
    }
    float ym[CX], yp[CX];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const float4 m = reg[((wm * 2 + 1) * H + h) * 64 + lane];
        const float4 p = reg[((wp * 2 + 0) * H + h) * 64 + lane];
        ym[4 * h] = m.x; ym[4 * h + 1] = m.y; ym[4 * h + 2] = m.z; ym[4 * h + 3] = m.w;
        yp[4 * h] = p.x; yp[4 * h + 1] = p.y; yp[4 * h + 2] = p.z; yp[4 * h + 3] = p.w;
    }
    // y neighbours per row as values (no select between arrays, which would
    // keep them in scratch memory)
    float ya[RPW][CX], yb[RPW][CX];
#pragma unroll
    for (int c = 0; c < CX; ++c) {
        ya[0][c] = ym[c];
        yb[RPW - 1][c] = yp[c];
    }
#pragma unroll
    for (int r = 1; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < CX; ++c) {
            ya[r][c] = nw.v[r - 1][c];
            yb[r - 1][c] = nw.v[r][c];
        }
    const bool first = lane == 0, last = lane == 63;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        // column 4*lane-1 of quad h is lane-1's last cell of quad h, except
        // on lane 0 where it is lane 63's last cell of quad h-1; likewise on
        // the right
        float pa[H], pb[H];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            pa[h] = from_prev_lane(nw.v[r][4 * h + 3]);
            pb[h] = from_next_lane(nw.v[r][4 * h]);
        }
        float xa[CX], xb[CX];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            if (h == 0)  // lane 0: outside the grid, or the seam (right half)
                xa[0] = (WX == 2 && first && wx == 1) ? sc[r] : pa[0];
            else
                xa[4 * h] = first ? pa[h - 1] : pa[h];
            if (h == H - 1)  // lane 63: outside the grid, or the seam (left half)
                xb[CX - 1] = (WX == 2 && last && wx == 0) ? sc[r] : pb[H - 1];
            else
                xb[4 * h + 3] = last ? pb[h + 1] : pb[h];
#pragma unroll
            for (int j = 1; j < 4; ++j) {
                xa[4 * h + j] = nw.v[r][4 * h + j - 1];
                xb[4 * h + j - 1] = nw.v[r][4 * h + j];
            }
        }
#pragma unroll
        for (int c = 0; c < CX; ++c) {
            float s = 0.0f;
            s += xa[c];
            s += xb[c];
            s += ya[r][c];
            s += yb[r][c];
            s += zm.v[r][c];
            out.v[r][c] = s;
        }
    }
}

// out = finish(partial + z-next) with cells outside the interior taking the
// centre value (equal shells: a shell cell never changes).  SH1 (level 1 of
// a fused pass with different shells): shell cells take u2's values instead
// (the shell plane k1 is loaded from S1 here, the x-shell values come from
// the LDS slice, the y-shell rows were loaded one step ahead into `sh`).
template <int OP, int CX, int RPW, bool SH1>
__device__ __forceinline__ void sm_finish(SmPlane<CX, RPW> &out, const SmPlane<CX, RPW> &part,
                                          const SmPlane<CX, RPW> &zp, const SmPlane<CX, RPW> &cen,
                                          float kc, bool zshell, const SmTile<RPW> &t,
                                          const SmLane<CX> &l, const SmShell1<CX, RPW> &sh,
                                          const SmXSlice &xsl, int w, const float *__restrict__ S1,
                                          const SmArgs &a, int k)
{
    typedef float V4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < CX; ++c) out.v[r][c] = finish<OP>(part.v[r][c] + zp.v[r][c], cen.v[r][c], kc);
    if (zshell) {
        if (SH1 && (k == 0 || k == a.nz - 1)) {
            // a z-shell plane: u2's values (twice per z column of tiles)
            const float *base = S1 + (long long)k * a.sz;
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int h = 0; h < CX / 4; ++h) {
                    const V4 q = *reinterpret_cast<const V4 *>(base + t.rowoff[r] + l.xo[h]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) out.v[r][4 * h + j] = q[j];
                }
            // wait for these loads inside this rare branch: the plane shares
            // its registers with the computed one, and a wait placed after
            // the join would hold every step until the prefetch of the next
            // plane has landed (vmcnt(0), expcnt and lgkmcnt untouched)
            __builtin_amdgcn_s_waitcnt(0x0F70);
        } else {
            out = cen;
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        if (t.yshell[r]) {
#pragma unroll
            for (int c = 0; c < CX; ++c) out.v[r][c] = (SH1 && t.ytrue[r]) ? sh.y1[r][c] : cen.v[r][c];
        } else {
            float2 x1 = make_float2(0.0f, 0.0f);
            if constexpr (SH1) x1 = xsl.at(1, w * RPW + r, k);
#pragma unroll
            for (int h = 0; h < CX / 4; ++h) {
                const float a0 = SH1 ? x1.x : cen.v[r][4 * h], b0 = out.v[r][4 * h];
                const float a1 = SH1 ? x1.y : cen.v[r][4 * h + 3], b1 = out.v[r][4 * h + 3];
                out.v[r][4 * h] = l.lo[h] ? a0 : b0;
                out.v[r][4 * h + 3] = l.hi[h] ? a1 : b1;
            }
        }
    }
}

// Store the interior rows of plane k.  With equal shells (SmLane::whole)
// the x shell cells, which hold their unchanged value, are stored too, so
// every row is written as whole cache lines (a partly written line costs
// HBM a read-modify-write); otherwise they are skipped.
template <int CX, int RPW>
__device__ __forceinline__ void sm_store(float *__restrict__ Y, const SmPlane<CX, RPW> &v,
                                         const SmArgs &a, const SmTile<RPW> &t, const SmLane<CX> &l,
                                         int k)
{
    if (k < t.z0 || k >= t.z1 || k < 1 || k > a.nz - 2) return;
    typedef float V4 __attribute__((ext_vector_type(4)));
    float *base = Y + (long long)k * a.sz;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        if (!t.ystore[r]) continue;
        float *row = base + t.rowoff[r];
#pragma unroll
        for (int h = 0; h < CX / 4; ++h) {
            if (l.whole[h]) {
                V4 q = {v.v[r][4 * h], v.v[r][4 * h + 1], v.v[r][4 * h + 2], v.v[r][4 * h + 3]};
                __builtin_nontemporal_store(q, reinterpret_cast<V4 *>(row + l.xo[h]));
            } else if (l.act[h]) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (!(j == 0 && l.lo[h]) && !(j == 3 && l.hi[h])) row[l.xo[h] + j] = v.v[r][4 * h + j];
            }
        }
    }
}

// One pass of T (1 or 2) sweeps X -> Y over all tiles of this workgroup.
// DIFF (T = 2, the buffers' shells differ): level-0 shell values from S0
// (u1), level-1 shell values from S1 (u2).
template <int OP, int CX, int RPW, int T, bool DIFF, int WX>
__device__ __forceinline__ void sm_pass(const float *__restrict__ X, float *__restrict__ Y,
                                        const float *__restrict__ S0, const float *__restrict__ S1,
                                        const float *__restrict__ xs, const SmArgs &a, float kc, float4 *lds,
                                        unsigned bid, unsigned nblk, bool same)
{
    static_assert(!DIFF || T == 2, "single-sweep passes read the shell of the buffer they stream");
    constexpr int H = CX / 4;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = blockDim.x >> 6;  // waves
    const int wy = nwv / WX;          // wave rows
    const int wr = w / WX;            // this wave's row of waves
    const int wx = w % WX;            // and its half of the row
    SmLane<CX> l;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int x = 256 * (wx * H + h) + 4 * lane;
        l.act[h] = x < a.nx;
        l.xo[h] = l.act[h] ? x : a.nx - 4;  // quads past the row reload its last cells
        l.lo[h] = x == 0;
        l.hi[h] = x + 4 == a.nx;
        l.whole[h] = l.act[h] && (same || (!l.lo[h] && !l.hi[h]));
    }
    float4 *reg0 = lds;
    float4 *reg1 = lds + 2 * nwv * H * 64;
    SmXSlice xsl;
    xsl.v = reinterpret_cast<float2 *>(lds + 4 * nwv * H * 64);
    xsl.rows = wy * RPW;
    xsl.stride = a.zchunk + 5;

    for (int tile = (int)bid; tile < a.ntiles; tile += (int)nblk) {
        const int tyi = tile % a.tiles_y, tzi = tile / a.tiles_y;
        SmTile<RPW> t;
        const int fy0 = tyi * a.core_h - 2;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            // wave-uniform row constants, kept in scalar registers
            const int fr = wr * RPW + r;  // frame row
            const int gy = __builtin_amdgcn_readfirstlane(fy0 + fr);
            const int gyc = min(max(gy, 0), a.ny - 1);
            t.rowoff[r] = gyc * a.nx;
            t.rowc[r] = gyc;
            t.yshell[r] = gy <= 0 || gy >= a.ny - 1;
            t.ytrue[r] = gy == 0 || gy == a.ny - 1;
            t.ystore[r] = !t.yshell[r] && fr >= 2 && fr < 2 + a.core_h;
        }
        t.z0 = tzi * a.zchunk;
        t.z1 = min(t.z0 + a.zchunk, a.nz);
        if constexpr (DIFF) {
            // the tile's x-shell values, planes z0-3 .. z1+1 (level 0 reads
            // z0-2 .. z1+1, level 1 z0-3 .. z1), both buffers: contiguous
            // runs of the table per row
            xsl.k0 = t.z0 - 3;
            const int nk = t.z1 - t.z0 + 5;
            const int n = 2 * xsl.rows * nk;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int row = i / nk;  // b * rows + frame row
                const int kk = i - row * nk;
                const int b = row >= xsl.rows ? 1 : 0;
                const int j = min(max(fy0 + row - b * xsl.rows, 0), a.ny - 1);
                const int k = min(max(xsl.k0 + kk, 0), a.nz - 1);
                xsl.v[row * xsl.stride + kk] =
                    *reinterpret_cast<const float2 *>(xs + xshell_index(b, k, j, 0, a.ny, a.nz));
            }
            __syncthreads();
        }

        // Pipeline state: the prefetched level-0 plane, the previous level-0
        // and level-1 planes, and the running partial sums P1/P2 (updated in
        // place once the finished plane has taken their old values).
        // (DIFF: sh1 holds the level-1 y-shell rows of the plane the step
        // finishes at level 1)
        SmPlane<CX, RPW> pf, c0, c1, P1, P2;
        SmShell1<CX, RPW> sh1;
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CX; ++c) c0.v[r][c] = c1.v[r][c] = P1.v[r][c] = P2.v[r][c] = 0.0f;
        const int kin_lo = t.z0 - T, kin_hi = t.z1 + T;
        sm_load<CX, RPW, DIFF>(pf, X, S0, a, t, l, kin_lo);
        if constexpr (DIFF) sm_load_sh1<CX, RPW>(sh1, S1, a, t, l, kin_lo - 1);
        for (int kin = kin_lo; kin < kin_hi; ++kin) {
            SmPlane<CX, RPW> nw0;
            sm_take<CX, RPW, DIFF>(nw0, pf, xsl, wr, kin, l);
            if (kin + 1 < kin_hi) sm_load<CX, RPW, DIFF>(pf, X, S0, a, t, l, kin + 1);
            // level 1, plane kin-1
            const int k1 = kin - 1;
            SmPlane<CX, RPW> o1;
            sm_finish<OP, CX, RPW, DIFF>(o1, P1, nw0, c0, kc, k1 <= 0 || k1 >= a.nz - 1, t, l, sh1, xsl, wr, S1,
                                         a, k1);
            // the level-1 y-shell rows the next step needs (plane kin)
            if constexpr (DIFF) sm_load_sh1<CX, RPW>(sh1, S1, a, t, l, kin);
            // one sweep: alternate the two regions so that a plane's LDS
            // writes never meet the previous plane's reads
            sm_partial<CX, RPW, WX>(P1, nw0, c0, (T == 1 && (kin & 1)) ? reg1 : reg0, w, wy, lane);
            c0 = nw0;
            if constexpr (T == 1) {
                sm_store<CX, RPW>(Y, o1, a, t, l, k1);
            } else {
                // level 2, plane kin-2
                const int k2 = kin - 2;
                SmPlane<CX, RPW> o2;
                sm_finish<OP, CX, RPW, false>(o2, P2, o1, c1, kc, k2 <= 0 || k2 >= a.nz - 1, t, l, sh1, xsl, wr,
                                              S1, a, k2);
                sm_partial<CX, RPW, WX>(P2, o1, c1, reg1, w, wy, lane);
                c1 = o1;
                sm_store<CX, RPW>(Y, o2, a, t, l, k2);
            }
        }
        __syncthreads();  // the next tile rewrites both LDS regions (and the slice)
    }
}

// Threads per workgroup from the state per lane (five planes of CX*RPW
// values plus the plane in flight): 8 values fit the 128 VGPRs of 16 waves,
// 16 the 168 of 12 waves, 32 the 256 of 8 waves.  The kernel for different
// shells (DIFF) also carries the shell values that travel with two of the
// planes: 8 values fit 12 waves, 16 and more 8 waves.  One row of 16 cells
// (per-row temporaries of 16 values) takes 8 waves in both kernels.
template <int CX, int RPW, bool DIFF>
constexpr int sm_max_threads()
{
    if (CX > 8) return 512;
    return DIFF ? (CX * RPW <= 8 ? 768 : 512) : (CX * RPW <= 8 ? 1024 : CX * RPW <= 16 ? 768 : 512);
}

// Two kernels share every launch of the streaming form: the one for equal
// shells (DIFF false) and the one for different shells (DIFF true); both
// read the comparison flags and the one whose case does not hold returns at
// once.  Single-sweep passes read the shell of the buffer they stream, which
// is right for either case, and run in whichever kernel stays.
template <int OP, int CX, int RPW, bool DIFF, int WX>
__global__ __launch_bounds__((sm_max_threads<CX, RPW, DIFF>())) void k_stream(float *__restrict__ u1,
                                                                            float *__restrict__ u2, SmArgs a,
                                                                            float kc)
{
    extern __shared__ float4 lds[];
    int ok = 1;
    for (int i = threadIdx.x; i < kShellBlocks; i += blockDim.x) ok &= a.shell_eq[i];
    const bool same = __syncthreads_and(ok) != 0;
    if (same == DIFF) return;
    const unsigned nblk = gridDim.x;
    const unsigned bid = xcd_remap(blockIdx.x, nblk);
    const int npass = min(a.n2 + a.n1, a.p1);
    for (int p = a.p0; p < npass; ++p) {
        const float *X = (p & 1) ? u2 : u1;
        float *Y = (p & 1) ? u1 : u2;
        if (p < a.n2) sm_pass<OP, CX, RPW, 2, DIFF, WX>(X, Y, u1, u2, a.xs, a, kc, lds, bid, nblk, same);
        else sm_pass<OP, CX, RPW, 1, false, WX>(X, Y, u1, u2, a.xs, a, kc, lds, bid, nblk, same);
        if (p + 1 < npass) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------- host side

// Cells per lane: 4 up to 256 columns, 8 up to 512 (one wave per row) and
// up to 1024 (two waves per row; coarsen_y = 1 asks for one wave of 16
// cells per lane instead).
static int stream_cx(const Cfg &c)
{
    if (c.nx % 4 || c.nx > 1024) return 0;
    return c.nx <= 256 ? 4 : 8;
}

// AUTO's use of the form: wherever it applies.  (With one wave of 16 cells
// per row over 512 columns, the row kernel was faster at 768^3, 647 against
// 601 GCells/s; with two waves of 8 per row the form gives 980 at 640^3,
// 1120 at 768^3 and 1263 at 896^3 against 650 for the row kernel.)
bool stream_auto_ok(const Cfg &c) { return stream_eligible(c); }

bool stream_eligible(const Cfg &c)
{
    return c.dtype == ST_F32 && c.dims == 3 && c.r == 1 && c.iters >= 1 &&
           c.nx >= 8 && c.ny >= 3 && c.nz >= 3 && stream_cx(c) != 0;
}

template <int CX, int RPW, bool DIFF, int WX = 1>
static const void *stream_fn_t(int op)
{
    return op == OP_THESIS ? reinterpret_cast<const void *>(&k_stream<OP_THESIS, CX, RPW, DIFF, WX>)
                           : reinterpret_cast<const void *>(&k_stream<OP_JACOBI, CX, RPW, DIFF, WX>);
}

// Instantiated shapes: 8 cells x 2 rows on two waves per row (nx up to
// 1024), 16 cells x 1 row (nx up to 1024), 8 cells x 2 rows (nx up to 512),
// 4 cells x 2 or 4 rows (nx up to 256).  8 x 4 needs more than 256 VGPRs
// and is not built.
static const void *stream_fn(int op, int cx, int rpw, bool diff, int wx)
{
    if (wx == 2) return diff ? stream_fn_t<8, 2, true, 2>(op) : stream_fn_t<8, 2, false, 2>(op);
    if (cx == 16) return diff ? stream_fn_t<16, 1, true>(op) : stream_fn_t<16, 1, false>(op);
    if (cx == 8) return diff ? stream_fn_t<8, 2, true>(op) : stream_fn_t<8, 2, false>(op);
    if (rpw == 4) return diff ? stream_fn_t<4, 4, true>(op) : stream_fn_t<4, 4, false>(op);
    return diff ? stream_fn_t<4, 2, true>(op) : stream_fn_t<4, 2, false>(op);
}

// Waves per frame that both kernels of a shape take.
static int stream_max_waves(int cx, int rpw)
{
    if (cx == 16) return std::min(sm_max_threads<16, 1, false>(), sm_max_threads<16, 1, true>()) / 64;
    if (cx == 8) return std::min(sm_max_threads<8, 2, false>(), sm_max_threads<8, 2, true>()) / 64;
    if (rpw == 4) return std::min(sm_max_threads<4, 4, false>(), sm_max_threads<4, 4, true>()) / 64;
    return std::min(sm_max_threads<4, 2, false>(), sm_max_threads<4, 2, true>()) / 64;
}

// LDS of a workgroup of wy wave rows of wx waves: the two exchange regions
// (2 edge rows per wave, in float4 of CX/4 quads per lane), then the
// tile's x-shell slice (2 buffers x frame rows x (chunk + 5) planes, float2
// each).
static int stream_lds(int wy, int rpw, int cx, int chunk, int wx)
{
    return 2 * (2 * wy * wx * (cx / 4) * 64 * 16) + 2 * wy * rpw * (chunk + 5) * 8;
}

// Dynamic LDS the form may take: the 160 KiB less the kernel's static LDS
// (256 bytes by the compiler's resource report; hipFuncSetAttribute refused
// a request of 163776 bytes).
constexpr int kStreamMaxLds = kMaxLds - 256;

// Deepest z chunk whose slice still fits the LDS.
static int stream_max_chunk(int wy, int rpw, int cx, int wx)
{
    return (kStreamMaxLds - stream_lds(wy, rpw, cx, -5, wx)) / (2 * wy * rpw * 8) - 5;
}

// Tiles of one pass for `wy` waves of `rpw` rows per frame: y frames times
// z chunks, the chunk count filling the CUs once (chunks no deeper than the
// LDS slice allows).
struct SmPlan {
    int wy, core_h, tiles_y, chunk, ntiles;
    double cost;
};

static SmPlan stream_plan(const Cfg &c, int wy, int rpw, int cx, int wx, int cus, int chunk_req)
{
    SmPlan p;
    p.wy = wy;
    p.core_h = wy * rpw - 4;
    p.tiles_y = (c.ny + p.core_h - 1) / p.core_h;
    int chunk = chunk_req;
    if (chunk <= 0) {
        int chunks = cus / p.tiles_y;
        if (chunks < 1) chunks = 1;
        chunk = (c.nz + chunks - 1) / chunks;
        if (chunk < 4) chunk = 4;
        if (chunk > c.nz) chunk = c.nz;
        const int cmax = stream_max_chunk(wy, rpw, cx, wx);
        if (chunk > cmax) {
            // the LDS caps the chunk: among the chunk counts from the
            // fewest that fit, the one whose busiest CU takes the least
            // time, as rounds of tiles x (planes per tile + 30: a tile's
            // start and its halo planes cost about 30 planes, as fitted to
            // 1024^3 with 128 frames of 8 core rows: 4 chunks of 256 (512
            // tiles, 2 rounds) 1423 GCells/s, 6 of 171 (3 rounds) 1365, 8
            // of 128 (4 rounds) 1273, 5 of 205 (2.5 rounds) 1269)
            const int n0 = (c.nz + cmax - 1) / cmax;
            long long best = -1;
            for (int n = n0; n < n0 + 8 && n <= c.nz; ++n) {
                const int ch = (c.nz + n - 1) / n;
                const long long nt = (long long)p.tiles_y * ((c.nz + ch - 1) / ch);
                const long long cost = (nt + cus - 1) / cus * (ch + 30LL);
                if (best < 0 || cost < best) {
                    best = cost;
                    chunk = ch;
                }
            }
        }
    }
    p.chunk = chunk;
    p.ntiles = p.tiles_y * ((c.nz + chunk - 1) / chunk);
    // time of the busiest CU: tile rounds x planes marched x HBM rows per
    // plane (core rows read and written; the four halo rows mostly hit L2)
    const int rounds = (p.ntiles + cus - 1) / cus;
    p.cost = (double)rounds * (chunk + 4) * (2.0 * p.core_h + 1.2);
    return p;
}

// The library's frame and chunk for one kernel (at most wmax wave rows).
static void stream_choose_auto(const Cfg &c, int rpw, int cx, int wx, int wmax, int cus, int chunk_req, int wmin,
                               SmPlan &best)
{
    // whole grid in one frame if it fits, else the frame height whose
    // tiles fill the CUs best
    const int wfit = (c.ny + 4 + rpw - 1) / rpw;
    best = stream_plan(c, wfit < wmax ? (wfit < wmin ? wmin : wfit) : wmax, rpw, cx, wx, cus, chunk_req);
    // candidate frame heights: down to two thirds of the largest, and
    // not below 8 waves where 8 fit (fewer waves lose the latency
    // hiding; 512^3: 6 waves x 3 chunks ran 21% slower than 8 x 5)
    const int wlo = std::max(std::max(wmin, (wmax * 2) / 3), std::min(wmax, 8));
    if (wfit > wmax)
        for (int wy = wmax - 1; wy >= wlo; --wy) {
            const SmPlan q = stream_plan(c, wy, rpw, cx, wx, cus, chunk_req);
            if (q.cost < best.cost * 0.97) best = q;  // fewer waves must pay for themselves
        }
    // Fields of 2.5 times the 256 MiB last-level cache and more: the
    // plan with the fewest planes marched per plane stored ((chunk + 4)
    // / chunk) wins as long as 75-92% of the CUs hold a tile (frames of
    // 2/3 of the largest height or more, as measured; ties go to more
    // tiles).  At 512^3: 8 waves x 5 chunks of 103 planes (215 tiles)
    // 9.67 ms per 100 sweeps, 10 x 8 chunks of 64 (256 tiles, the
    // busiest-CU model's choice) 10.2, 12 x 8 of 64 (208) 9.88, with the
    // same HBM bytes; at 448^3 8 x 6 of 75 (228 tiles) 6.72 against 6.98.
    // At 384^3 the busiest-CU model's 256 tiles stay the fastest.
    if (chunk_req <= 0 && 2LL * c.nx * c.ny * c.nz * 4 >= (640LL << 20)) {
        SmPlan pick{};
        double pick_amp = 0;
        for (int wy = wmax; wy >= wlo; --wy) {
            const int core_h = wy * rpw - 4;
            const int tiles_y = (c.ny + core_h - 1) / core_h;
            for (int nch = 1; nch * tiles_y <= cus; ++nch) {
                const int chunk = (c.nz + nch - 1) / nch;
                const int ntiles = tiles_y * ((c.nz + chunk - 1) / chunk);
                if (chunk < 4 || chunk > stream_max_chunk(wy, rpw, cx, wx) || ntiles * 100 < cus * 75 ||
                    ntiles * 100 > cus * 92)
                    continue;
                const double amp = (chunk + 4.0) / chunk;
                if (pick.wy == 0 || amp < pick_amp - 1e-9 || (amp < pick_amp + 1e-9 && ntiles > pick.ntiles)) {
                    pick = stream_plan(c, wy, rpw, cx, wx, cus, chunk);
                    pick_amp = amp;
                }
            }
        }
        if (pick.wy) best = pick;
    }
}

// Frame and chunk of one kernel of the form (at most wmax wave rows).
static st_status stream_choose(const Cfg &c, int rpw, int cx, int wx, int wmax, int cus, SmPlan &best)
{
    const int chunk_req = c.zchunk > 0 ? c.zchunk : 0;
    const int wmin = (4 + rpw) / rpw;  // at least one core row
    if (!c.heuristic && c.by > 0) {
        // block_y = wave rows per frame (frame of rpw*block_y rows)
        if (c.by < wmin || c.by > wmax)
            return fail(ST_EINVAL, "streaming form: block_y (wave rows per frame) must be %d..%d", wmin, wmax);
        best = stream_plan(c, c.by, rpw, cx, wx, cus, chunk_req);
    } else {
        stream_choose_auto(c, rpw, cx, wx, wmax, cus, chunk_req, wmin, best);
    }
    const int lds = stream_lds(best.wy, rpw, cx, best.chunk, wx);
    if (lds > kStreamMaxLds)
        return fail(ST_ELDS, "streaming form: a z chunk of %d planes needs %d bytes of LDS (at most %d "
                             "planes with this frame)", best.chunk, lds, stream_max_chunk(best.wy, rpw, cx, wx));
    return ST_OK;
}

st_status resolve_stream(Cfg &c)
{
    if (!stream_eligible(c))
        return fail(ST_EUNSUPPORTED, "streaming temporal form needs FP32, 3D, radius 1, nx a multiple "
                                     "of 4 up to 1024");
    int cx = stream_cx(c);
    // rows per wave: 2, or coarsen_y = 4 with 4 cells per lane (nx <= 256);
    // over 512 columns two waves per row, or coarsen_y = 1: one row of 16
    // cells per lane
    int rpw = 2, wx = c.nx > 512 ? 2 : 1;
    if (!c.heuristic && c.cy > 0) {
        if (c.cy != 2 && !(c.cy == 4 && cx == 4) && !(c.cy == 1 && c.nx > 512))
            return fail(ST_EINVAL, "streaming form: coarsen_y must be 2 (or 4 when nx <= 256; 1 when "
                                   "nx > 512)");
        rpw = c.cy;
        if (rpw == 1) {
            cx = 16;
            wx = 1;
        }
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    // Wave rows per frame at most: both kernels (equal and different
    // shells) take the frame of the one with fewer registers to spare,
    // except with two waves per row, where the kernel for equal shells
    // takes 12 waves (6 wave rows, 8 core rows of 12) and the other 8 (4
    // of 8): frames of 8 rows recompute every stored row once more.
    const int wmax_d = stream_max_waves(cx, rpw) / wx;
    const int wmax = wx == 2 ? sm_max_threads<8, 2, false>() / 64 / wx : wmax_d;
    SmPlan best, bestd;
    st_status st = stream_choose(c, rpw, cx, wx, wmax, cus, best);
    if (st != ST_OK) return st;
    if (wmax != wmax_d) {
        Cfg cd = c;
        if (!cd.heuristic && cd.by > wmax_d) cd.by = wmax_d;
        st = stream_choose(cd, rpw, cx, wx, wmax_d, cus, bestd);
        if (st != ST_OK) return st;
    } else {
        bestd = best;
    }
    c.kind = 3;
    c.cx = cx;
    c.cy = rpw;
    c.wx = wx;
    c.wy = best.wy;
    c.tsteps = 2;
    c.lds = stream_lds(best.wy, rpw, cx, best.chunk, wx);
    c.bx = c.nx;
    c.by = best.core_h;
    c.zchunk = best.chunk;
    c.bz = best.chunk;
    c.wy_d = bestd.wy;
    c.core_d = bestd.core_h;
    c.zchunk_d = bestd.chunk;
    c.lds_d = stream_lds(bestd.wy, rpw, cx, bestd.chunk, wx);
    c.gx = 0;
    c.gy = c.gz = 1;
    return ST_OK;
}

st_status run_stream(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2)
{
    *in_u2 = c.iters & 1;
    if (c.iters <= 0) return ST_OK;
    // workspace: the shell-comparison flags, then the x-shell table of both
    // buffers (2 x 2 values per row)
    const size_t flag_bytes = 1024;
    static_assert(kShellBlocks * sizeof(int) <= flag_bytes, "flags fit their part");
    const size_t xs_bytes = (size_t)2 * 2 * c.ny * c.nz * sizeof(float);
    void *ws = nullptr;
    st_status st = workspace_slot(3, flag_bytes + xs_bytes, s, &ws);
    if (st != ST_OK) return st;
    float *xs = reinterpret_cast<float *>(static_cast<unsigned char *>(ws) + flag_bytes);
    hipLaunchKernelGGL(k_shell_eq, dim3(kShellBlocks), dim3(256), 0, s, static_cast<const float *>(u1),
                       static_cast<const float *>(u2), c.nx, c.ny, c.nz, static_cast<int *>(ws), xs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_check(e, "shell comparison launch");

    const void *fns[2] = {stream_fn(c.op, c.cx, c.cy, false, c.wx), stream_fn(c.op, c.cx, c.cy, true, c.wx)};
    // each kernel's frame: [0] equal shells, [1] different shells
    SmArgs a[2];
    int threads[2], grid[2], lds[2];
    for (int k = 0; k < 2; ++k) {
        const int wy = k ? c.wy_d : c.wy, chunk = k ? c.zchunk_d : c.zchunk;
        SmArgs &b = a[k];
        b.nx = c.nx; b.ny = c.ny; b.nz = c.nz;
        b.sz = (long long)c.nx * c.ny;
        b.core_h = wy * c.cy - 4;
        b.tiles_y = (c.ny + b.core_h - 1) / b.core_h;
        b.zchunk = chunk;
        b.ntiles = b.tiles_y * ((c.nz + chunk - 1) / chunk);
        b.iters = c.iters;
        b.n1 = c.iters % 4;
        b.n2 = (c.iters - b.n1) / 2;
        b.shell_eq = static_cast<const int *>(ws);
        b.xs = xs;
        threads[k] = 64 * wy * c.wx;
        lds[k] = k ? c.lds_d : c.lds;
        e = hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, lds[k]);
        if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(stream LDS)");
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fns[k], threads[k], lds[k]);
        if (e != hipSuccess) return hip_check(e, "occupancy query");
        if (per_cu < 1) return fail(ST_ELDS, "streaming kernel does not fit on a CU");
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid[k] = std::min(b.ntiles, per_cu * cus);
    }
    float kf = c.op == OP_THESIS ? (float)(2 * c.dims * c.r) : (float)c.weight;
    void *args0[] = {&u1, &u2, &a[0], &kf};
    void *args1[] = {&u1, &u2, &a[1], &kf};
    void **args[2] = {args0, args1};
    // Which kernel runs: outside graph capture (and outside the no-alloc
    // entry point) the host reads the comparison flags once and launches
    // only the kernel for the case that holds; otherwise both are launched
    // and the device decides (the other returns at once, a few
    // microseconds per launch).
    int k_lo = 0, k_hi = 2;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!workspace_is_fixed() && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
        int flags[kShellBlocks];
        e = hipMemcpyAsync(flags, ws, sizeof flags, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_check(e, "shell comparison read-back");
        int same = 1;
        for (int i = 0; i < kShellBlocks; ++i) same &= flags[i] != 0;
        k_lo = same ? 0 : 1;
        k_hi = k_lo + 1;
    }
    if (!c.per_pass) {
        // one cooperative launch per kernel, a grid barrier between passes
        bool coop = true;
        for (int k = k_lo; k < k_hi && coop; ++k) {
            a[k].p0 = 0;
            a[k].p1 = a[k].n2 + a[k].n1;
            e = hipLaunchCooperativeKernel(fns[k], dim3(grid[k]), dim3(threads[k]), args[k], lds[k], s);
            if (e == hipErrorIllegalState && k == k_lo) {
                // the device takes no cooperative launch from this process
                // now: the same passes follow, one launch per pass
                (void)hipGetLastError();
                coop = false;
            } else if (e != hipSuccess) {
                return hip_check(e, "hipLaunchCooperativeKernel(stream)");
            }
        }
        if (coop) return ST_OK;
    }
    // one launch per pass and kernel, kernel boundaries as the barrier: n2
    // fused passes, then n1 single-sweep passes (the kernel for different
    // shells runs them too, when that is the case)
    for (int p = 0; p < a[0].n2 + a[0].n1; ++p) {
        for (int k = k_lo; k < k_hi; ++k) {
            a[k].p0 = p;
            a[k].p1 = p + 1;
            e = hipLaunchKernel(fns[k], dim3(grid[k]), dim3(threads[k]), args[k], lds[k], s);
            if (e != hipSuccess) return hip_check(e, "stream pass launch");
        }
    }
    return ST_OK;
}

}  // namespace st
