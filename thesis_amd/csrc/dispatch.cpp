// Host side of the C ABI: knob validation, launch-shape resolution, the
// sweep loop and small HIP utilities.
//
// Replaces src/stencil_dispatch.cu:14-158 (get_kernel, set_smem,
// set_max_occupancy_block_dimensions, set_block_dims,
// set_max_dynamic_shared_memory_size, dispatch_kernels) and the validation
// of include/stencil_utils.h:113-124 (check_early_exit).  Errors are returned
// as status codes; nothing here exits the process.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>

#include "st_internal.hpp"

namespace st {

static thread_local std::string g_last_error;

st_status fail(st_status s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

st_status hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return ST_OK;
    if (e == hipErrorOutOfMemory) return fail(ST_ENOMEM, "%s: %s", what, hipGetErrorString(e));
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        return fail(ST_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    return fail(ST_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------- workspace

// One workspace per (device, stream, slot): calls on different streams never
// share device scratch (shell flags, reduction partials, batch tables), and
// calls on one stream are ordered by the stream.
static std::mutex g_ws_mutex;
constexpr int kSlots = 4;
struct WsBlock {
    void *ptr = nullptr;
    size_t size = 0;
};
static std::map<std::tuple<int, hipStream_t, int>, WsBlock> g_ws;

// While set on this thread, workspace requests are served from a fixed
// device buffer (one part per slot) and nothing is allocated.
static thread_local const WsFixed *g_ws_fixed = nullptr;

void workspace_use_fixed(const WsFixed *fixed) { g_ws_fixed = fixed; }
bool workspace_is_fixed() { return g_ws_fixed != nullptr; }

st_status workspace_slot(int slot, size_t bytes, hipStream_t stream, void **ptr)
{
    if (slot < 0 || slot >= kSlots) return fail(ST_EINVAL, "workspace slot %d", slot);
    if (g_ws_fixed) {
        const size_t part = g_ws_fixed->bytes[slot];
        if (bytes > part)
            return fail(ST_EINVAL, "workspace of %zu bytes exceeds the fixed scratch part of %zu bytes (slot %d)",
                        bytes, part, slot);
        *ptr = static_cast<unsigned char *>(g_ws_fixed->base) + g_ws_fixed->offset[slot];
        return ST_OK;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_check(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    WsBlock &w = g_ws[std::make_tuple(dev, stream, slot)];
    if (w.size < bytes) {
        if (w.ptr) {
            // earlier work on this stream may still use the old block
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_check(e, "workspace sync");
            (void)hipFree(w.ptr);
            w.ptr = nullptr;
            w.size = 0;
        }
        size_t sz = bytes < (1u << 20) ? (1u << 20) : bytes;
        e = hipMalloc(&w.ptr, sz);
        if (e != hipSuccess) return hip_check(e, "workspace hipMalloc");
        w.size = sz;
    }
    *ptr = w.ptr;
    return ST_OK;
}

// ---------------------------------------------------------------- validation

static bool dtype_sweepable(int dt) { return dt == ST_F32 || dt == ST_F64; }

static st_status validate_common(const st_params *p, bool mixed)
{
    if (!p) return fail(ST_EINVAL, "null params");
    if (p->dimensions != 2 && p->dimensions != 3)
        return fail(ST_EINVAL, "dimensions must be 2 or 3 (got %d)", p->dimensions);
    if (p->radius < 1) return fail(ST_EINVAL, "radius must be >= 1 (got %d)", p->radius);
    if (p->iterations < 0) return fail(ST_EINVAL, "iterations must be >= 0");
    const int nz = p->dimensions == 3 ? p->nz : 1;
    if (p->nx <= 2 * p->radius || p->ny <= 2 * p->radius ||
        (p->dimensions == 3 && nz <= 2 * p->radius))
        return fail(ST_EINVAL, "extents (%d,%d,%d) must exceed 2*radius=%d", p->nx, p->ny, nz,
                    2 * p->radius);
    if ((long long)p->nx * p->ny >= (1LL << 31))
        return fail(ST_EINVAL, "nx*ny must be below 2^31");
    if ((long long)p->nx * p->ny * nz >= (1LL << 40)) return fail(ST_EINVAL, "grid too large");
    if (p->op != ST_OP_THESIS && p->op != ST_OP_JACOBI) return fail(ST_EINVAL, "unknown operator");
    if (p->variant < ST_V_BASE || p->variant > ST_V_AUTO) return fail(ST_EINVAL, "unknown variant");
    if (mixed) {
        if (p->dtype != ST_BF16 && p->dtype != ST_F16)
            return fail(ST_EINVAL, "mixed precision needs dtype BF16 or F16");
        if (p->op != ST_OP_JACOBI) return fail(ST_EINVAL, "mixed precision uses the JACOBI operator");
    } else if (!dtype_sweepable(p->dtype)) {
        return fail(ST_EINVAL, "sweeps store F32 or F64; BF16/F16 go through stencil_dispatch_mixed");
    }
    if (p->coarsen_x < 0 || p->coarsen_y < 0 || p->smem_pad < 0 || p->temporal_steps < 0 ||
        p->launch_per_pass < 0 || p->launch_per_pass > 1 || p->temporal_form < 0 ||
        p->temporal_form > 5)
        return fail(ST_EINVAL, "coarsening, padding, temporal steps, launch_per_pass or "
                               "temporal_form out of range");
    const bool heur = p->heuristic || p->block_x <= 0;
    if (!heur) {
        if (p->block_x < 1 || p->block_y < 1 || (p->dimensions == 3 && p->block_z < 1))
            return fail(ST_EINVAL, "block dims must be positive");
        const long long thr = (long long)p->block_x * p->block_y *
                              (p->dimensions == 3 && p->variant <= ST_V_SMEM_PADDED ? p->block_z : 1);
        if (p->variant <= ST_V_SMEM_PADDED && thr > kMaxThreads)
            return fail(ST_EINVAL, "%lld threads per workgroup exceed 1024", thr);
        // check_early_exit (include/stencil_utils.h:113-124)
        if (p->variant == ST_V_SMEM_PADDED) {
            if (p->block_x < p->radius || p->block_y < p->radius)
                return fail(ST_EINVAL, "SMEM_PADDED needs block_x, block_y >= radius");
            if (p->dimensions == 3 && p->block_z < p->radius)
                return fail(ST_EINVAL, "SMEM_PADDED needs block_z >= radius");
        }
        if (p->variant == ST_V_SMEM_REGISTER) {
            if (p->block_x < p->radius) return fail(ST_EINVAL, "SMEM_REGISTER needs block_x >= radius");
            if (p->dimensions == 3 && p->block_y < p->radius)
                return fail(ST_EINVAL, "SMEM_REGISTER needs block_y >= radius");
        }
    }
    if (p->weight < 0 || std::isnan(p->weight)) return fail(ST_EINVAL, "weight must be >= 0");
    return ST_OK;
}

static st_status resolve(const st_params *p, Cfg &c, bool mixed)
{
    st_status s = validate_common(p, mixed);
    if (s != ST_OK) return s;
    c.nx = p->nx; c.ny = p->ny; c.nz = p->dimensions == 3 ? p->nz : 1;
    c.dims = p->dimensions; c.r = p->radius; c.iters = p->iterations;
    c.op = p->op; c.dtype = p->dtype; c.variant = p->variant;
    const bool heur = p->heuristic || p->block_x <= 0;
    c.bx = heur ? 0 : p->block_x;
    c.by = heur ? 0 : p->block_y;
    c.bz = heur ? 0 : (p->dimensions == 3 ? p->block_z : 1);
    c.cx = p->coarsen_x > 0 ? p->coarsen_x : 1;
    c.cy = p->coarsen_y;
    c.pad = p->smem_pad;
    c.tsteps = p->temporal_steps;
    c.weight = p->weight > 0 ? p->weight : 1.0 / (2.0 * c.dims * c.r);
    c.weight64 = c.weight;
    if (c.dtype == ST_F32 || c.dtype == ST_BF16 || c.dtype == ST_F16) c.weight = (double)(float)c.weight;
    c.restart = p->restart > 0 ? p->restart : 10;
    c.heuristic = p->heuristic;
    c.per_pass = p->launch_per_pass;
    c.form = p->temporal_form;
    c.kind = 0;
    c.wx = c.wy = 0;
    c.hs = 1;
    c.rowc = 0;
    c.wy_d = c.core_d = c.zchunk_d = c.lds_d = 0;
    c.shell_eq = nullptr;
    c.zchunk = 0;
    c.lds = 0;
    c.gx = c.gy = c.gz = 1;
    if (mixed) return ST_OK;
    if (c.variant == ST_V_AUTO) {
        // Measured on MI355X (FP32 R=1, warm, per sweep):
        //   * the streaming temporal form runs one launch per pass: a
        //     grid-wide barrier of a cooperative launch costs 25-30 us on
        //     this device, a kernel boundary a few (512^3, 100 sweeps: 10.1 ms
        //     against 11.4 for one cooperative launch);
        //   * it wins over the per-sweep row kernel from 320^3 up, with equal
        //     shells (two sweeps per pass; 320^3: 1065 against 771 GCells/s)
        //     and with different shells (one sweep per pass; 622 against
        //     567); at 288^3 and below the row kernel, served from the
        //     256 MiB last-level cache, wins with different shells (288^3:
        //     660 against 554) though not with equal ones (739 against 935),
        //     and the shells are compared on the device only, so the switch
        //     sits between 288^3 and 320^3 (both fields above 192 MiB);
        //   * grids of up to about a million cells (3D) or four million (2D)
        //     are bound by the launch cadence, where the plain kernel is the
        //     cheapest (64^3: 2.5 us against 5.2 for the row kernel and 8.3
        //     for the streaming form; 256^2: 2.5 against 7.5 for the 2D
        //     register march);
        //   * elsewhere the register march / row kernel wherever it is
        //     instantiated, else the plain kernel.
        c.bx = c.by = c.bz = 0;
        c.heuristic = 1;
        const long long cells = (long long)c.nx * c.ny * c.nz;
        const long long es = c.dtype == ST_F64 ? 8 : 4;
        const long long stream_bytes = 192LL << 20;
        const long long small = c.dims == 3 ? (1LL << 20) : (1LL << 22);
        if (c.iters >= 4 && (c.tsteps == 0 || c.tsteps == 2) && c.form == 0 && stream_auto_ok(c) &&
            2 * cells * es > stream_bytes) {
            c.variant = ST_V_COOP;
            c.cx = 1;
            c.cy = 0;
            c.per_pass = 1;
        } else if (c.dims == 2 && c.iters >= 8 && c.form == 0 && (c.tsteps == 0 || c.tsteps == 4 || c.tsteps == 6) &&
                   stream2d_eligible(c) && cells >= small) {
            // 2D beyond the launch-bound sizes: six or four fused sweeps per
            // pass (64 sweeps: 4096^2 1511 GCells/s against 771 for the
            // register march, 2048^2 570 against 404 for the plain kernel;
            // 1024^2 stays with the plain kernel, 212 against 143)
            c.variant = ST_V_COOP;
            c.cx = 1;
            c.cy = 0;
            c.per_pass = 1;
        } else if (cells <= small) {
            c.variant = ST_V_BASE;
            c.cx = 1;
        } else {
            const bool march_ok = (c.dtype == ST_F32 && c.r <= 8) ||
                                  (c.dtype == ST_F64 && (c.r == 1 || c.r == 2 || c.r == 4));
            c.variant = march_ok ? ST_V_SMEM_REGISTER : ST_V_BASE;
            if (!march_ok) c.cx = 1;
        }
    }
    switch (c.variant) {
    case ST_V_BASE: return resolve_base(c);
    case ST_V_SMEM: return resolve_smem(c);
    case ST_V_SMEM_PADDED: return resolve_padded(c);
    case ST_V_SMEM_REGISTER: return resolve_march(c);
    case ST_V_COOP: return resolve_coop(c);
    default: return fail(ST_EINVAL, "unknown variant");
    }
}

static st_status sweep_once(const Cfg &c, const void *in, void *out, hipStream_t s)
{
    switch (c.variant) {
    case ST_V_BASE: return launch_base(c, in, out, s);
    case ST_V_SMEM: return launch_smem(c, in, out, s);
    case ST_V_SMEM_PADDED: return launch_padded(c, in, out, s);
    case ST_V_SMEM_REGISTER: return launch_march(c, in, out, s);
    default: return fail(ST_EUNSUPPORTED, "variant %d has no single-sweep launch", c.variant);
    }
}


}  // namespace st

using namespace st;

extern "C" {

void stencil_params_default(st_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->nx = p->ny = p->nz = 256;  // DIM
    p->dimensions = 3;
    p->radius = 1;
    p->iterations = 64;
    p->block_x = 32;
    p->block_y = 2;
    p->block_z = 2;
    p->coarsen_x = 4;
    p->variant = ST_V_BASE;
    p->op = ST_OP_THESIS;
    p->dtype = ST_F32;
    p->heuristic = 0;
}

st_status stencil_validate(const st_params *p)
{
    Cfg c;
    return resolve(p, c, false);
}

st_status stencil_launch_config(const st_params *p, st_launch_info *info)
{
    Cfg c;
    st_status s = resolve(p, c, false);
    if (s != ST_OK) return s;
    if (info) {
        info->variant = c.variant;
        info->block_x = c.bx; info->block_y = c.by; info->block_z = c.bz;
        info->grid_x = c.gx; info->grid_y = c.gy; info->grid_z = c.gz;
        info->coarsen_x = c.cx; info->coarsen_y = c.cy;
        info->zchunk = c.zchunk;
        info->temporal_steps = c.tsteps;
        info->lds_bytes = c.lds;
        if (c.variant == ST_V_COOP) {
            int n = c.iters > 0 ? (c.iters + c.tsteps - 1) / c.tsteps : 0;
            if (n && (n & 1) != (c.iters & 1)) ++n;
            if (c.kind == 2) n = c.iters / c.tsteps;
            // streaming form per pass: n2 fused passes, then iters mod 4
            // single ones
            if (c.kind == 3) n = c.iters > 0 ? (c.per_pass ? (c.iters - c.iters % 4) / 2 + c.iters % 4 : 1) : 0;
            if (c.kind == 5 || c.kind == 7) n = c.iters;
            info->launches = n ? (c.per_pass ? n : 1) : 0;
            // 2D streaming form, always one launch per pass (with equal
            // shells: fused passes of six and four sweeps, then single sweeps)
            if (c.kind == 6) info->launches = stream2d_launches(c);
        } else {
            info->launches = c.iters;
        }
    }
    return ST_OK;
}

st_status stencil_sweep(const st_params *p, const void *d_in, void *d_out, void *stream)
{
    Cfg c;
    st_status s = resolve(p, c, false);
    if (s != ST_OK) return s;
    if (!d_in || !d_out || d_in == d_out) return fail(ST_EINVAL, "sweep needs two distinct buffers");
    return sweep_once(c, d_in, d_out, static_cast<hipStream_t>(stream));
}

st_status stencil_dispatch(const st_params *p, void *d_u1, void *d_u2, void *stream,
                           int *result_in_u2, float *elapsed_ms)
{
    Cfg c;
    st_status s = resolve(p, c, false);
    if (s != ST_OK) return s;
    if (!d_u1 || !d_u2 || d_u1 == d_u2) return fail(ST_EINVAL, "dispatch needs two distinct buffers");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        if ((s = hip_check(hipEventCreate(&e0), "event create")) != ST_OK) return s;
        if ((s = hip_check(hipEventCreate(&e1), "event create")) != ST_OK) {
            (void)hipEventDestroy(e0);
            return s;
        }
        if ((s = hip_check(hipEventRecord(e0, hs), "event record")) != ST_OK) goto done;
    }
    {
        int in_u2 = 0;
        if (c.variant == ST_V_COOP) {
            s = run_coop(c, d_u1, d_u2, hs, &in_u2);
        } else {
            // the row form stores whole rows when both shells are equal
            // (a shell cell then holds the same bits at every sweep)
            if (c.variant == ST_V_SMEM_REGISTER && c.kind == 4 && c.iters > 1)
                s = rows_shell_check(c, d_u1, d_u2, hs, &c.shell_eq);
            void *a = d_u1, *b = d_u2;
            for (int i = 0; i < c.iters && s == ST_OK; ++i) {
                s = sweep_once(c, a, b, hs);
                void *t = a; a = b; b = t;
            }
            in_u2 = (a == d_u2);
        }
        if (result_in_u2) *result_in_u2 = in_u2;
    }
    if (s == ST_OK && elapsed_ms) {
        if ((s = hip_check(hipEventRecord(e1, hs), "event record")) != ST_OK) goto done;
        if ((s = hip_check(hipEventSynchronize(e1), "event sync")) != ST_OK) goto done;
        s = hip_check(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed time");
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return s;
}

// Initialisation and the residual depend on the geometry only: resolve them
// with the kernel-shape knobs reset so that any sweep configuration works.
static st_status resolve_geometry(const st_params *p, Cfg &c)
{
    if (!p) return fail(ST_EINVAL, "null params");
    st_params q = *p;
    q.variant = ST_V_BASE;
    q.heuristic = 1;
    q.block_x = q.block_y = q.block_z = 0;
    q.coarsen_x = 1;
    q.coarsen_y = 0;
    return resolve(&q, c, false);
}

st_status stencil_init(const st_params *p, int kind, uint64_t seed, void *d_u1, void *d_u2,
                       void *stream)
{
    Cfg c;
    st_status s = resolve_geometry(p, c);
    if (s != ST_OK) return s;
    if (kind != ST_INIT_REFERENCE && kind != ST_INIT_SEEDED) return fail(ST_EINVAL, "unknown init kind");
    if (!d_u1 || !d_u2) return fail(ST_EINVAL, "null buffer");
    return launch_init(c, kind, seed, d_u1, d_u2, static_cast<hipStream_t>(stream));
}

st_status stencil_residual_norm(const st_params *p, const void *d_u, double *norm, void *stream)
{
    Cfg c;
    st_status s = resolve_geometry(p, c);
    if (s != ST_OK) return s;
    if (!d_u || !norm) return fail(ST_EINVAL, "null argument");
    return residual_norm(c, d_u, norm, static_cast<hipStream_t>(stream));
}

st_status stencil_dispatch_mixed(const st_params *p, void *d_x, void *d_e1, void *d_e2, void *d_d,
                                 void *d_u, void *stream, float *elapsed_ms)
{
    Cfg c;
    st_status s = resolve(p, c, true);
    if (s != ST_OK) return s;
    if (!d_x || !d_e1 || !d_e2 || !d_d || !d_u) return fail(ST_EINVAL, "null buffer");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, hs);
    }
    s = run_mixed(c, d_x, d_e1, d_e2, d_d, d_u, hs);
    if (elapsed_ms) {
        if (s == ST_OK) {
            (void)hipEventRecord(e1, hs);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(elapsed_ms, e0, e1);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return s;
}

st_status stencil_dispatch_batched(const st_problem *problems, int nproblems,
                                   const st_tile_rule *rules, int nrules, int mode, void *stream,
                                   float *elapsed_ms)
{
    if (mode < ST_BATCH_GRAPH || mode > ST_BATCH_SEQUENTIAL) return fail(ST_EINVAL, "unknown batch mode");
    if (nrules < 0 || (nrules > 0 && !rules)) return fail(ST_EINVAL, "bad rule table");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    hipStream_t own = nullptr;
    if (!hs && mode == ST_BATCH_GRAPH) {
        st_status s = hip_check(hipStreamCreate(&own), "stream create");
        if (s != ST_OK) return s;
        hs = own;
    }
    st_status s = run_batched(problems, nproblems, rules, nrules, mode, hs, elapsed_ms);
    if (own) {
        (void)hipStreamSynchronize(own);
        (void)hipStreamDestroy(own);
    }
    return s;
}

const char *stencil_status_str(st_status s)
{
    switch (s) {
    case ST_OK: return "ok";
    case ST_EINVAL: return "invalid argument";
    case ST_ELDS: return "LDS budget exceeded";
    case ST_ELAUNCH: return "launch failed";
    case ST_ENOMEM: return "out of memory";
    case ST_EUNSUPPORTED: return "unsupported configuration";
    case ST_EDEVICE: return "no usable device";
    }
    return "unknown status";
}

const char *stencil_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------- host utilities

st_status st_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return n > 0 ? ST_OK : fail(ST_EDEVICE, "no HIP device");
}

st_status st_device_name(int device, char *buf, int len)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_check(e, "hipGetDeviceProperties");
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return ST_OK;
}

st_status st_malloc(void **ptr, uint64_t bytes) { return hip_check(hipMalloc(ptr, bytes), "hipMalloc"); }
st_status st_free(void *ptr) { return hip_check(hipFree(ptr), "hipFree"); }

st_status st_memcpy_htod(void *dst, const void *src, uint64_t bytes, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    st_status r = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "memcpy htod");
    return r == ST_OK ? hip_check(hipStreamSynchronize(s), "sync") : r;
}

st_status st_memcpy_dtoh(void *dst, const void *src, uint64_t bytes, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    st_status r = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "memcpy dtoh");
    return r == ST_OK ? hip_check(hipStreamSynchronize(s), "sync") : r;
}

st_status st_memcpy_dtod(void *dst, const void *src, uint64_t bytes, void *stream)
{
    return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                                    static_cast<hipStream_t>(stream)), "memcpy dtod");
}

st_status st_memset(void *dst, int value, uint64_t bytes, void *stream)
{
    return hip_check(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)), "memset");
}

st_status st_stream_create(void **stream)
{
    hipStream_t s = nullptr;
    st_status r = hip_check(hipStreamCreate(&s), "stream create");
    if (stream) *stream = s;
    return r;
}

st_status st_stream_destroy(void *stream)
{
    return hip_check(hipStreamDestroy(static_cast<hipStream_t>(stream)), "stream destroy");
}

st_status st_stream_sync(void *stream)
{
    return hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "stream sync");
}

st_status st_device_sync(void) { return hip_check(hipDeviceSynchronize(), "device sync"); }

}  // extern "C"
