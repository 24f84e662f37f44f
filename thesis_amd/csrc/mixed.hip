// 16-bit field storage (BF16 or FP16) with FP32 arithmetic and FP32 defect
// correction, for the JACOBI operator.  No reference counterpart; the
// arithmetic is restated bit for bit in oracle/stencil_oracle.c
// (or_mixed_run) and checked against it.
//
// The field is u = x + e: x an FP32 base field that carries the Dirichlet
// shell, e a 16-bit correction and d a 16-bit defect d = G(x) - x.  Because
// the sweep G is affine, G(x + e) = G(x) + G0(e), so one sweep of u is the
// correction sweep  e' = d + G0(e)  which moves 2 + 2 + 2 bytes per cell.
// Every `restart` sweeps the correction is folded into x in FP32 and the
// defect is recomputed in FP32 from the new x (the FP32-residual defect
// correction); 16-bit rounding then only ever applies to increments.
#include "st_device.hpp"
#include "st_half.hpp"
#include "st_internal.hpp"

namespace st {

struct MixArgs {
    Geom g;
    int dims;
    float w;
};

// Decompose interior index t into (i, j, k).
__device__ __forceinline__ long long interior_idx(const MixArgs &a, long long t)
{
    const Geom &g = a.g;
    const int R = g.r;
    const int wx = g.nx - 2 * R, wy = g.ny - 2 * R;
    const long long wxy = (long long)wx * wy;
    const int kk = (int)(t / wxy);
    const long long rem = t - (long long)kk * wxy;
    const int j = R + (int)(rem / wx);
    const int i = R + (int)(rem - (long long)(j - R) * wx);
    const int k = a.dims == 3 ? R + kk : 0;
    return i + j * g.sy + k * g.sz;
}

__device__ __forceinline__ long long interior_count(const MixArgs &a)
{
    const Geom &g = a.g;
    const int R = g.r;
    return (long long)(g.nx - 2 * R) * (g.ny - 2 * R) * (a.dims == 3 ? g.nz - 2 * R : 1);
}

// d = round16(G(x) - x) on the interior; G is the FP32 JACOBI cell.
template <bool F16>
__global__ __launch_bounds__(256) void k_mx_defect(const float *__restrict__ x,
                                                   uint16_t *__restrict__ d, MixArgs a)
{
    const Geom &g = a.g;
    const int R = g.r;
    const long long n = interior_count(a), stride = (long long)gridDim.x * blockDim.x;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += stride) {
        const long long idx = interior_idx(a, t);
        float u = 0.0f;
        for (int q = R; q >= 1; --q) u += x[idx - q];
        for (int q = 1; q <= R; ++q) u += x[idx + q];
        for (int q = R; q >= 1; --q) u += x[idx - q * g.sy];
        for (int q = 1; q <= R; ++q) u += x[idx + q * g.sy];
        if (a.dims == 3) {
            for (int q = R; q >= 1; --q) u += x[idx - q * g.sz];
            for (int q = 1; q <= R; ++q) u += x[idx + q * g.sz];
        }
        const float tcell = u * a.w;
        d[idx] = rnd16<F16>(tcell - x[idx]);
    }
}

// e' = round16(widen(d) + w * sum widen(e neighbours)).
template <bool F16>
__global__ __launch_bounds__(256) void k_mx_sweep(const uint16_t *__restrict__ e,
                                                  const uint16_t *__restrict__ d,
                                                  uint16_t *__restrict__ eo, MixArgs a)
{
    const Geom &g = a.g;
    const int R = g.r;
    const long long n = interior_count(a), stride = (long long)gridDim.x * blockDim.x;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += stride) {
        const long long idx = interior_idx(a, t);
        float s = 0.0f;
        for (int q = R; q >= 1; --q) s += wid16<F16>(e[idx - q]);
        for (int q = 1; q <= R; ++q) s += wid16<F16>(e[idx + q]);
        for (int q = R; q >= 1; --q) s += wid16<F16>(e[idx - q * g.sy]);
        for (int q = 1; q <= R; ++q) s += wid16<F16>(e[idx + q * g.sy]);
        if (a.dims == 3) {
            for (int q = R; q >= 1; --q) s += wid16<F16>(e[idx - q * g.sz]);
            for (int q = 1; q <= R; ++q) s += wid16<F16>(e[idx + q * g.sz]);
        }
        const float sw = s * a.w;
        eo[idx] = rnd16<F16>(wid16<F16>(d[idx]) + sw);
    }
}

// v <- 0 over n 16-bit values (a kernel rather than a memset, so that the
// sequence can be captured in a graph and replayed)
__global__ void k_fill16(uint16_t *__restrict__ v, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) v[i] = 0;
}

// x += widen(e); e <- 0 (whole grid, elementwise).
template <bool F16>
__global__ __launch_bounds__(256) void k_mx_fold(float *__restrict__ x, uint16_t *__restrict__ e,
                                                 long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i] = x[i] + wid16<F16>(e[i]);
        e[i] = 0;
    }
}

// u = x + widen(e).
template <bool F16>
__global__ __launch_bounds__(256) void k_mx_final(const float *__restrict__ x,
                                                  const uint16_t *__restrict__ e,
                                                  float *__restrict__ u, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
        u[i] = x[i] + wid16<F16>(e[i]);
}

// u += widen(e) (the final fold when the current field is already in u).
template <bool F16>
__global__ __launch_bounds__(256) void k_mx_final_inplace(float *__restrict__ u, const uint16_t *__restrict__ e,
                                                          long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
        u[i] = u[i] + wid16<F16>(e[i]);
}

template <bool F16>
static st_status mixed_impl(const Cfg &c, float *x, uint16_t *e1, uint16_t *e2, uint16_t *d,
                            float *u, hipStream_t s)
{
    MixArgs a;
    a.g.nx = c.nx; a.g.ny = c.ny; a.g.nz = c.nz; a.g.r = c.r;
    a.g.sy = c.nx; a.g.sz = (long long)c.nx * c.ny;
    a.dims = c.dims;
    a.w = (float)c.weight;
    const long long n = (long long)c.nx * c.ny * c.nz;
    const int R = c.r;
    const long long ni = (long long)(c.nx - 2 * R) * (c.ny - 2 * R) * (c.dims == 3 ? c.nz - 2 * R : 1);
    auto blocks = [](long long cnt) {
        long long b = (cnt + 255) / 256;
        if (b > 32768) b = 32768;
        return dim3((unsigned)(b < 1 ? 1 : b));
    };
    hipLaunchKernelGGL(k_fill16, blocks(n), dim3(256), 0, s, e1, n);
    hipLaunchKernelGGL(k_fill16, blocks(n), dim3(256), 0, s, e2, n);
    hipLaunchKernelGGL(k_fill16, blocks(n), dim3(256), 0, s, d, n);
    // 3D grids the row-streaming kernels take (FP32 arithmetic on 16-bit
    // loads, same operation order); the rest use the per-cell kernels
    const bool rows = c.dims == 3 && rows_mixed_eligible(c);
    st_status st = ST_OK;
    if (rows && c.iters > 0 && rows_mixed_fused_ok(c)) {
        // The same arithmetic with two passes fewer per restart: the defect
        // sweep also writes the first correction sweep of its cycle
        // (round16(widen(d) + w*0), which is what a sweep from a zero
        // correction computes), and the fold x += e of a restart is formed
        // on the fly by the next defect sweep, which writes x' to the other
        // FP32 buffer (u, given x's shell first) instead of updating x in
        // place.  Sweep it (1-based) is the first of its cycle when
        // (it - 1) % restart == 0.
        if (hipMemcpyAsync(u, x, n * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
            return hip_check(hipGetLastError(), "mixed: copy of x");
        float *xa = x, *xb = u;
        uint16_t *ea = e1, *eb = e2;
        st = launch_rows_defect(c, F16, xa, nullptr, nullptr, d, ea, s);
        for (int it = 2; it <= c.iters && st == ST_OK; ++it) {
            if ((it - 1) % c.restart == 0) {
                st = launch_rows_defect(c, F16, xa, ea, xb, d, eb, s);
                float *t = xa; xa = xb; xb = t;
            } else {
                st = launch_rows_mixed(c, 0, F16, ea, eb, d, s);
            }
            uint16_t *t = ea; ea = eb; eb = t;
        }
        if (st != ST_OK) return st;
        if (xa == u) hipLaunchKernelGGL(k_mx_final_inplace<F16>, blocks(n), dim3(256), 0, s, u, ea, n);
        else hipLaunchKernelGGL(k_mx_final<F16>, blocks(n), dim3(256), 0, s, xa, ea, u, n);
        return hip_check(hipGetLastError(), "mixed launch");
    }
    auto defect = [&]() {
        if (rows) st = launch_rows_mixed(c, 1, F16, x, d, nullptr, s);
        else hipLaunchKernelGGL(k_mx_defect<F16>, blocks(ni), dim3(256), 0, s, x, d, a);
    };
    auto sweep = [&](const uint16_t *ein, uint16_t *eout) {
        if (rows) st = launch_rows_mixed(c, 0, F16, ein, eout, d, s);
        else hipLaunchKernelGGL(k_mx_sweep<F16>, blocks(ni), dim3(256), 0, s, ein, d, eout, a);
    };
    defect();
    uint16_t *ea = e1, *eb = e2;
    for (int it = 1; it <= c.iters && st == ST_OK; ++it) {
        sweep(ea, eb);
        uint16_t *t = ea; ea = eb; eb = t;
        if (it % c.restart == 0 && it < c.iters) {
            // eb's interior is overwritten by the next sweep and its shell is
            // still 0, so only the live correction needs clearing.
            hipLaunchKernelGGL(k_mx_fold<F16>, blocks(n), dim3(256), 0, s, x, ea, n);
            defect();
        }
    }
    if (st != ST_OK) return st;
    hipLaunchKernelGGL(k_mx_final<F16>, blocks(n), dim3(256), 0, s, x, ea, u, n);
    return hip_check(hipGetLastError(), "mixed launch");
}

st_status run_mixed(const Cfg &c, void *x, void *e1, void *e2, void *d, void *u, hipStream_t s)
{
    if (c.dtype == ST_F16)
        return mixed_impl<true>(c, static_cast<float *>(x), static_cast<uint16_t *>(e1),
                                static_cast<uint16_t *>(e2), static_cast<uint16_t *>(d),
                                static_cast<float *>(u), s);
    return mixed_impl<false>(c, static_cast<float *>(x), static_cast<uint16_t *>(e1),
                             static_cast<uint16_t *>(e2), static_cast<uint16_t *>(d),
                             static_cast<float *>(u), s);
}

}  // namespace st

namespace st {

template <bool F16>
__global__ void k_round16(const float *__restrict__ in, uint16_t *__restrict__ out, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = rnd16<F16>(in[i]);
}

template <bool F16>
__global__ void k_widen16(const uint16_t *__restrict__ in, float *__restrict__ out, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = wid16<F16>(in[i]);
}

static dim3 conv_blocks(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b < 1 ? 1 : b));
}

}  // namespace st

extern "C" st_status stencil_round16(const float *d_in, uint16_t *d_out, uint64_t n, int dtype, void *stream)
{
    using namespace st;
    if (dtype != ST_BF16 && dtype != ST_F16) return fail(ST_EINVAL, "round16: dtype must be BF16 or F16");
    if (!d_in || !d_out) return fail(ST_EINVAL, "round16: null buffer");
    if (n == 0) return ST_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == ST_F16) hipLaunchKernelGGL(k_round16<true>, conv_blocks(n), dim3(256), 0, s, d_in, d_out, (long long)n);
    else hipLaunchKernelGGL(k_round16<false>, conv_blocks(n), dim3(256), 0, s, d_in, d_out, (long long)n);
    return hip_check(hipGetLastError(), "round16 launch");
}

extern "C" st_status stencil_widen16(const uint16_t *d_in, float *d_out, uint64_t n, int dtype, void *stream)
{
    using namespace st;
    if (dtype != ST_BF16 && dtype != ST_F16) return fail(ST_EINVAL, "widen16: dtype must be BF16 or F16");
    if (!d_in || !d_out) return fail(ST_EINVAL, "widen16: null buffer");
    if (n == 0) return ST_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == ST_F16) hipLaunchKernelGGL(k_widen16<true>, conv_blocks(n), dim3(256), 0, s, d_in, d_out, (long long)n);
    else hipLaunchKernelGGL(k_widen16<false>, conv_blocks(n), dim3(256), 0, s, d_in, d_out, (long long)n);
    return hip_check(hipGetLastError(), "widen16 launch");
}
