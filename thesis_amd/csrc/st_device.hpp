// Device-side helpers shared by the gfx950 sweep kernels.
//
// Arithmetic contract (bit-exact with the reference's stencil(),
// src/stencils.cu:8-19, and with oracle/stencil_oracle.c):
//   sum = 0; sum += i-prev (d = R..1); sum += i-next (d = 1..R);
//   j-prev; j-next; k-prev; k-next   (each one IEEE add, in this order)
//   THESIS: out = (sum / coeff) - centre      coeff = (T)(2*D*R)
//   JACOBI: out = sum * w
// Every file is compiled with -ffp-contract=off so no add/multiply pair is
// fused, and float division stays correctly rounded (HIP default).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

enum : int { OP_THESIS = 0, OP_JACOBI = 1 };

// Grid geometry.  nz == 1 for 2D.  Strides in elements.
struct Geom {
    int nx, ny, nz;
    int r;            // radius (runtime copy; kernels with a compile-time
                      // radius ignore it)
    long long sy;     // nx
    long long sz;     // nx*ny
};

template <int OP, typename T>
__device__ __forceinline__ T finish(T sum, T centre, T kc)
{
    if constexpr (OP == OP_THESIS) {
        T q = sum / kc;
        return q - centre;
    } else {
        return sum * kc;
    }
}

__device__ __forceinline__ bool interior1(int i, int n, int r) { return i >= r && i < n - r; }

// Cross-lane read of a 32-bit or 64-bit value from lane `src` (0..63; any
// other value wraps modulo 64) through ds_bpermute, which needs no LDS
// allocation and no barrier.
__device__ __forceinline__ float lane_read(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute((src & 63) << 2, __float_as_int(v)));
}

__device__ __forceinline__ double lane_read(double v, int src)
{
    long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    lo = __builtin_amdgcn_ds_bpermute((src & 63) << 2, lo);
    hi = __builtin_amdgcn_ds_bpermute((src & 63) << 2, hi);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Remap a linear workgroup id so that the workgroups the hardware places on
// one XCD (ids congruent modulo 8, dispatched round-robin) get a contiguous
// range of logical tiles; neighbouring tiles then share that XCD's L2.
// A bijection on [0, n).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n)
{
    const unsigned n8 = n & ~7u;
    if (b >= n8) return b;
    return (b & 7u) * (n8 >> 3) + (b >> 3);
}

// The R-thick shell of an nx x ny x nz grid (every cell that is not
// interior) as a list: first the full shell rows (the 2R outer planes, then
// the 2R outer rows of each inner plane) cell by cell along x, then the 2R
// end cells (x < R, x >= nx - R) of each inner row.  Consecutive list
// positions are consecutive memory in the first part, so lanes read it
// coalesced.
__host__ __device__ inline long long shell_full_cells(int nx, int ny, int nz, int R)
{
    return (2LL * R * ny + (long long)(nz - 2 * R) * 2 * R) * nx;
}
__host__ __device__ inline long long shell_cells(int nx, int ny, int nz, int R)
{
    const long long inner = (long long)(nz - 2 * R) * (ny - 2 * R);
    return shell_full_cells(nx, ny, nz, R) + (inner > 0 ? inner * 2 * R : 0);
}
// Linear index (x fastest) of list position i < shell_cells(...) (which the
// callers keep below 2^31, so the list arithmetic is 32-bit); n1 =
// shell_full_cells(...).
__device__ __forceinline__ long long shell_cell(unsigned i, unsigned n1, int nx, int ny, int nz, int R)
{
    unsigned x, y, z;
    const unsigned R2 = 2u * R;
    if (i < n1) {
        const unsigned row = i / nx;
        x = i - row * nx;
        if (row < R2 * ny) {       // an outer plane
            const unsigned pl = row / ny;
            z = pl < (unsigned)R ? pl : nz - R2 + pl;
            y = row - pl * ny;
        } else {                   // an outer row of an inner plane
            const unsigned r2 = row - R2 * ny;
            z = R + r2 / R2;
            const unsigned m = r2 % R2;
            y = m < (unsigned)R ? m : ny - R2 + m;
        }
    } else {
        const unsigned j = i - n1;
        const unsigned row = j / R2;
        const unsigned m = j - row * R2;
        const unsigned in = ny - R2;
        z = R + row / in;
        y = R + row % in;
        x = m < (unsigned)R ? m : nx - R2 + m;
    }
    return (long long)z * nx * ny + (long long)y * nx + x;
}

constexpr int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

}  // namespace st
