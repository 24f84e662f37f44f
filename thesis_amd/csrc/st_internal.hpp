// Library-internal host declarations: the resolved launch configuration and
// the per-family launchers implemented in the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/stencil_abi.h"

namespace st {

// Knobs after validation and heuristics (the runtime form of the
// reference's compile-time constants, include/constants.h:1-78).
struct Cfg {
    int nx, ny, nz;      // nz == 1 in 2D
    int dims, r, iters;
    int op, dtype, variant;
    int bx, by, bz;      // workgroup shape as launched
    int cx, cy;          // coarsening along x (cells per lane) and rows per lane
    int pad;             // LDS row padding (floats)
    int zchunk;          // planes per workgroup (register march)
    int tsteps;          // sweeps per pass (cooperative temporal blocking)
    int gx, gy, gz;      // grid
    int lds;             // dynamic LDS bytes
    double weight;       // JACOBI weight (resolved, rounded to the storage type)
    double weight64;     // the same weight unrounded (FP64 residual norm)
    int restart;         // mixed precision defect-correction interval
    int heuristic;       // HEURISTIC knob
    int per_pass;        // COOP: one launch per pass instead of a cooperative launch
    int form;            // COOP form requested (0 auto, 1 LDS ring, 2 register queues)
    int kind;            // COOP form resolved
    int wx, wy;          // COOP register form: waves across x and y
    int hs;              // SMEM_REGISTER row form: halo strips per wave
    int rowc;            // row form: 1 = consecutive columns per lane (k_rowc)
    int wy_d, core_d;    // COOP streaming form, kernel for different shells (when its
    int zchunk_d, lds_d; // frame differs): wave rows, core rows, z chunk, LDS bytes
    const int *shell_eq; // device flag: both buffers' shells hold the same bits
                         // (row form stores whole rows then); null: unknown
};

// Record a detail string for stencil_last_error() and return `s`.
st_status fail(st_status s, const char *fmt, ...);

st_status hip_check(hipError_t e, const char *what);

// Per-family launch-shape resolution (fills block/grid/lds in `c`).
st_status resolve_base(Cfg &c);
st_status resolve_smem(Cfg &c);
st_status resolve_padded(Cfg &c);
st_status resolve_march(Cfg &c);
st_status resolve_coop(Cfg &c);

// One sweep in -> out.
st_status launch_base(const Cfg &c, const void *in, void *out, hipStream_t s);
st_status launch_smem(const Cfg &c, const void *in, void *out, hipStream_t s);
st_status launch_padded(const Cfg &c, const void *in, void *out, hipStream_t s);
st_status launch_march(const Cfg &c, const void *in, void *out, hipStream_t s);

// Cooperative temporal-blocked run of c.iters sweeps; *in_u2 receives the
// buffer that holds the result.
st_status run_coop(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2);

// Register-queue temporal blocking (sweep_tblock.hip).
bool tblock_eligible(const Cfg &c, int t);
st_status resolve_tblock(Cfg &c, int t);
st_status run_tblock(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2);

// SMEM_REGISTER row-streaming form, one sweep (sweep_rows.hip).
bool rows_eligible(const Cfg &c);
st_status resolve_rows(Cfg &c);
st_status resolve_rows_coop(Cfg &c);
st_status run_rows_coop(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2);
st_status launch_rows(const Cfg &c, const void *in, void *out, hipStream_t s);
// 16-bit storage modes on the geometry of `c` (FP32 arithmetic, JACOBI
// weight c.weight): defect d = round16(w*sum(x) - x), else the correction
// e' = round16(widen(d) + w*sum(widen(e))); the output's shell must be zero
// (whole edge quads are stored, with zero in the shell cells)
bool rows_mixed_eligible(const Cfg &c);
// Compare the R-thick shells of u1 and u2 on the stream; *flag receives a
// device pointer to an int that is 1 when they hold the same bits.
st_status rows_shell_check(const Cfg &c, const void *u1, const void *u2, hipStream_t s, const int **flag);
st_status launch_rows_mixed(const Cfg &c, int defect, bool f16, const void *in, void *out, const uint16_t *d,
                            hipStream_t s);
// The consecutive row layout serves the correction, defect and fused fold
// modes on the geometry of `c`.
bool rows_mixed_fused_ok(const Cfg &c);
// Defect sweep d = round16(w*sum(x') - x') that also writes e_first =
// round16(widen(d) + 0), the first correction sweep from a zero correction
// (shell zero).  With e_old, x' = x + widen(e_old) is formed on the fly and
// its interior rows stored to x_out (a buffer other than x, holding x's
// shell); without, x' = x and x_out is unused.
st_status launch_rows_defect(const Cfg &c, bool f16, const float *x, const uint16_t *e_old, float *x_out,
                             uint16_t *d, uint16_t *e_first, hipStream_t s);

// 2D streaming temporal blocking, four or six fused sweeps per pass
// (sweep_stream2d.hip); `single` is the per-sweep configuration for the
// sweeps outside fused passes.
bool stream2d_eligible(const Cfg &c);
st_status resolve_stream2d(Cfg &c);
st_status run_stream2d(const Cfg &c, const Cfg &single, void *u1, void *u2, hipStream_t s, int *in_u2);
int stream2d_launches(const Cfg &c);  // launches with equal shells

// Streaming temporal blocking, whole rows per wave (sweep_stream.hip).
bool stream_eligible(const Cfg &c);
bool stream_auto_ok(const Cfg &c);  // AUTO takes the form for this shape
st_status resolve_stream(Cfg &c);
st_status run_stream(const Cfg &c, void *u1, void *u2, hipStream_t s, int *in_u2);

st_status launch_init(const Cfg &c, int kind, unsigned long long seed, void *u1, void *u2,
                      hipStream_t s);
st_status residual_norm(const Cfg &c, const void *u, double *norm, hipStream_t s);
st_status run_mixed(const Cfg &c, void *x, void *e1, void *e2, void *d, void *u,
                    hipStream_t s);

// Library-owned device workspace (grows on demand, one per device, stream
// and slot).
st_status workspace_slot(int slot, size_t bytes, hipStream_t stream, void **ptr);
// A fixed device buffer that serves every workspace request of the calling
// thread while it is set (one part per slot; a request larger than its
// slot's part fails with ST_EINVAL); null returns to the allocated
// workspace.
struct WsFixed {
    void *base;
    size_t offset[4];  // part of each slot: [offset, offset + bytes)
    size_t bytes[4];
};
void workspace_use_fixed(const WsFixed *fixed);
// True while a fixed buffer serves this thread's workspace (the no-alloc
// entry point, whose callers may be capturing a graph).
bool workspace_is_fixed();
inline st_status workspace(size_t bytes, hipStream_t stream, void **ptr) { return workspace_slot(0, bytes, stream, ptr); }

st_status run_batched(const st_problem *probs, int n, const st_tile_rule *rules, int nrules,
                      int mode, hipStream_t s, float *elapsed_ms);

constexpr int kMaxLds = 160 * 1024;
constexpr int kMaxThreads = 1024;

}  // namespace st
