/*
 * stencil_abi.h - C ABI of the MI355X (gfx950) iterative star-stencil sweep.
 *
 * This is the drop-in boundary for the hot path of mrfarstad/thesis:
 *   host  main()            -> dispatch_kernels(d_u1, d_u2)
 *                              src/main.cu:64-67, src/stencil_dispatch.cu:138-158
 *   launch  get_kernel()<<<grid, block, smem>>>(d_u1, d_u2, kstart, kend)
 *                              src/stencil_dispatch.cu:14-42,147-157
 *   knobs   -D DIM DIMENSIONS RADIUS ITERATIONS BLOCK_X/Y/Z COARSEN_X SMEM
 *              PADDED REGISTER COOP HEURISTIC SMEM_PAD
 *                              Makefile:21-71, include/constants.h:1-78
 * The compile-time knobs become fields of st_params; the kernel choice of
 * get_kernel() becomes st_params.variant.  All functions return a status
 * code instead of exiting the process (include/stencil_utils.h:113-124 and
 * include/helper_cuda.h:891-930 in the reference abort).
 *
 * Conventions
 *   - Fields are single contiguous arrays, x fastest:
 *       idx = i + j*nx + k*nx*ny          (src/stencil_kernel_base.cu:14)
 *   - Device pointers are plain HIP device addresses; `stream` is a
 *     hipStream_t passed as void* (NULL = the null stream).
 *   - Only interior cells (R <= i < n-R on every axis) change.  The
 *     R-thick boundary shell of both buffers keeps its bits for the whole
 *     run; when both buffers hold the same shell (compared on the GPU at
 *     the start of a dispatch) some kernels store shell cells of interior
 *     rows with the value they already hold, so rows are written as whole
 *     cache lines.
 *   - The caller owns every buffer.  Entry points use a small device
 *     workspace owned by the library (shell-comparison flags, reduction
 *     partials, the batch table), allocated on first use per device and
 *     stream and kept until the process ends.
 *   - Calls are stream-ordered.  Calls on different streams never share
 *     workspace; they are re-entrant for disjoint buffers and not
 *     thread-safe on the same buffers.
 */
#ifndef THESIS_AMD_STENCIL_ABI_H
#define THESIS_AMD_STENCIL_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum st_status {
    ST_OK = 0,
    ST_EINVAL = 1,        /* invalid knobs or shapes (check_early_exit analogue) */
    ST_ELDS = 2,          /* LDS request above the 160 KiB per workgroup limit   */
    ST_ELAUNCH = 3,       /* a kernel launch or HIP call failed                  */
    ST_ENOMEM = 4,        /* workspace allocation failed                         */
    ST_EUNSUPPORTED = 5,  /* valid knobs that this build does not implement      */
    ST_EDEVICE = 6        /* no usable gfx950 device                             */
} st_status;

/* Operator.  THESIS is the reference update (src/stencils.cu:8-19):
 *     u' = (sum of the 2*D*R axis neighbours) / (float)(2*D*R) - u_centre
 * JACOBI is the classical weighted-neighbour Jacobi sweep:
 *     u' = weight * (sum of the 2*D*R axis neighbours)
 * Both sum the neighbours in the reference order: i-prev (d=R..1),
 * i-next (d=1..R), j-prev, j-next, k-prev, k-next
 * (src/stencil_accumulators.cu:6-16,46-74). */
typedef enum st_op { ST_OP_THESIS = 0, ST_OP_JACOBI = 1 } st_op;

/* Kernel family, replacing get_kernel() (src/stencil_dispatch.cu:14-42):
 *   BASE          base_{2d,3d}, base_unroll_*      all operands from HBM/cache
 *   SMEM          smem_*, smem_unroll_*            LDS tile, halo from global
 *   SMEM_PADDED   smem_padded_*                    LDS tile with R halo
 *   SMEM_REGISTER smem_register_*                  register-queue march along
 *                                                  the slowest axis (z in 3D,
 *                                                  y in 2D), cross-lane x
 *   COOP          coop (src/stencil_kernel_coop.cu) persistent cooperative
 *                                                  kernel, grid-synchronised,
 *                                                  temporal blocking (radius
 *                                                  1..4; grid-stride sweeps
 *                                                  beyond)
 *   AUTO          fastest family for the shape (library's own choice)      */
typedef enum st_variant {
    ST_V_BASE = 0,
    ST_V_SMEM = 1,
    ST_V_SMEM_PADDED = 2,
    ST_V_SMEM_REGISTER = 3,
    ST_V_COOP = 4,
    ST_V_AUTO = 5
} st_variant;

/* Storage type of the fields.  F32 and F64 compute in the storage type.
 * BF16 and F16 store 16-bit fields and compute in FP32; they are used only by
 * stencil_dispatch_mixed. */
typedef enum st_dtype { ST_F32 = 0, ST_F64 = 1, ST_BF16 = 2, ST_F16 = 3 } st_dtype;

/* Initial conditions written by stencil_init. */
typedef enum st_init {
    /* initialize_host_region (include/stencil_initializer.h:6-52): the
     * outermost layer of u1 is 1.0, every other cell of u1 is 0; u2 is zeroed
     * (the reference leaves it uninitialised; fresh memory is zero). */
    ST_INIT_REFERENCE = 0,
    /* every cell of the R-thick boundary shell of both buffers gets
     * U(seed, idx) = (splitmix64(seed ^ idx*0x9E3779B97F4A7C15) >> 40) * 2^-24,
     * the interior of both buffers is 0 */
    ST_INIT_SEEDED = 1
} st_init;

typedef struct st_params {
    int nx, ny, nz;       /* extents; the reference's DIM gives nx=ny=nz; nz is
                             ignored (taken as 1) when dimensions == 2        */
    int dimensions;       /* DIMENSIONS: 2 or 3                               */
    int radius;           /* RADIUS >= 1                                       */
    int iterations;       /* ITERATIONS >= 0                                   */
    int block_x;          /* BLOCK_X/Y/Z: workgroup shape (0 = heuristic).     */
    int block_y;          /*   SMEM_REGISTER: block_x is a multiple of 64 (one */
    int block_z;          /*   wave across x), block_z is the z-chunk depth in */
                          /*   planes (0 = whole axis split heuristically).    */
                          /*   COOP: block_z is ignored, except by the         */
                          /*   streaming form (temporal_form 3), where block_y */
                          /*   is the waves per frame (2 rows each) and        */
                          /*   block_z the z-chunk depth.                      */
    int coarsen_x;        /* COARSEN_X: cells per thread along x (>= 1)        */
    int coarsen_y;        /* SMEM_REGISTER/COOP: rows per thread (0=heuristic) */
    int smem_pad;         /* SMEM_PAD: extra LDS floats per tile row           */
    int variant;          /* st_variant                                        */
    int op;               /* st_op                                             */
    int dtype;            /* st_dtype                                          */
    int heuristic;        /* HEURISTIC: 1 = library picks block/coarsening     */
    int temporal_steps;   /* COOP: sweeps fused per HBM pass (0 = heuristic)   */
    int restart;          /* mixed precision: sweeps between defect
                             corrections (0 = 10)                              */
    int launch_per_pass;  /* COOP: 0 = one persistent cooperative launch with
                             grid-wide barriers between passes, 1 = one launch
                             per pass (kernel boundaries as the barrier)       */
    int temporal_form;    /* COOP: 0 = library's choice, 1 = LDS-ring tiles
                             (F32/F64, radius 1..4, any pass length),
                             2 = register-queue full-row frames (F32, radius
                             1, temporal_steps 2, even iteration count),
                             3 = streaming rows, one wave per grid row up
                             to 512 columns and two above (F32, radius 1,
                             nx a multiple of 4 up to 1024; two fused
                             sweeps per pass whether or not the two buffers
                             hold the same shell; block_y = wave rows per
                             frame, coarsen_y = rows per wave, 1 over 512
                             columns asks for one wave of 16 cells per
                             lane); in 2D
                             (with the grid-stride form below) 512-column segments
                             per wave, six and four fused sweeps per pass
                             when both buffers hold the same shell (else
                             single sweeps; temporal_steps 4 keeps passes of
                             four), nx a multiple of 8, block_y = rows per
                             chunk,
                             4 = row frames of the SMEM_REGISTER row form, one
                             sweep per grid barrier (3D F32/F64, radius 1..4,
                             nx a multiple of 4),
                             5 = grid-stride cells, one sweep per grid
                             barrier (the reference's coop kernel,
                             src/stencil_kernel_coop.cu:5-49: 2D/3D,
                             F32/F64, any radius; the library's choice for
                             radius 5 and up and for 2D grids the streaming
                             form does not take).
                             SMEM_REGISTER: 0 = library's choice, 1 = register
                             march, 2 = row streaming (3D F32/F64, radius
                             1..4, nx a multiple of 4; block_y = rows per
                             frame, block_z = z-chunk depth)                */
    double weight;        /* JACOBI neighbour weight (0 = 1/(2*D*R)); rounded
                             to the storage type                               */
} st_params;

/* Resolved launch configuration (what get_kernel/set_block_dims/set_smem
 * compute in src/stencil_dispatch.cu:46-122). */
typedef struct st_launch_info {
    int variant;                 /* resolved st_variant (AUTO replaced)        */
    int block_x, block_y, block_z;
    int grid_x, grid_y, grid_z;
    int coarsen_x, coarsen_y;
    int zchunk;                  /* planes per workgroup (SMEM_REGISTER)       */
    int temporal_steps;          /* sweeps per pass (COOP)                     */
    int lds_bytes;               /* LDS per workgroup                          */
    int launches;                /* sweep-kernel launches for `iterations`
                                    sweeps (the shell comparison and the
                                    streaming form's band fix-ups are extra) */
} st_launch_info;

/* Default knobs of include/constants.h:1-52: DIM 256, DIMENSIONS 3,
 * RADIUS 1, ITERATIONS 64, BLOCK 32x2x2, COARSEN_X 4, base kernel,
 * THESIS operator, FP32. */
void stencil_params_default(st_params *p);

/* Host-only validation, the analogue of check_early_exit
 * (include/stencil_utils.h:113-124): SMEM_PADDED rejects block_x, block_y
 * (and block_z in 3D) below RADIUS; SMEM_REGISTER rejects block_x (and
 * block_y in 3D) below RADIUS.  Also rejects non-positive extents, extents not
 * larger than 2*RADIUS, more than 1024 threads, and LDS above 160 KiB. */
st_status stencil_validate(const st_params *p);

/* Resolve the launch configuration without launching anything. */
st_status stencil_launch_config(const st_params *p, st_launch_info *info);

/* One sweep d_out <- S(d_in): one get_kernel()<<<...>>> launch
 * (src/stencil_dispatch.cu:154).  d_in and d_out must not alias.
 * Not available for the COOP variant (which fuses sweeps). */
st_status stencil_sweep(const st_params *p, const void *d_in, void *d_out, void *stream);

/* dispatch_kernels (src/stencil_dispatch.cu:138-158): p->iterations sweeps
 * ping-ponging between d_u1 and d_u2.  On return *result_in_u2 (if not NULL)
 * is 1 when the final field is in d_u2 and 0 when it is in d_u1; as in the
 * reference it is d_u1 for an even iteration count.  The COOP variant
 * schedules its fused passes so that the result lands in the same buffer as
 * the per-sweep variants; the interior of the other buffer then holds an
 * intermediate state, not sweep iterations-1.  If elapsed_ms is not NULL the
 * call brackets the sweep loop with HIP events on `stream`, waits for the
 * stop event and stores the time in milliseconds (host set-up excluded). */
st_status stencil_dispatch(const st_params *p, void *d_u1, void *d_u2, void *stream,
                           int *result_in_u2, float *elapsed_ms);

/* stencil_dispatch without any allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  The same
 * sweeps and result as stencil_dispatch (elapsed time is not measured).  The
 * dispatch's device scratch comes from a static device array of 16 MiB +
 * 5 KiB (one per device, held until the process ends) instead of the
 * library's allocated workspace: the row form's shell flag (4 KiB part) and
 * the streaming form's shell flags and x-shell table (1 KiB + 16 B per grid
 * row: grids with ny*nz up to 2^20); a larger request fails with ST_EINVAL.
 * Calls that overlap on one device share that array and race; calls ordered
 * on one stream, or one after another, do not. */
st_status stencil_dispatch_noalloc(const st_params *p, void *d_u1, void *d_u2, void *stream,
                                   int *result_in_u2);

/* Write the initial condition `kind` (st_init) into d_u1 and d_u2 (both
 * nx*ny*nz elements of p->dtype, F32 or F64). */
st_status stencil_init(const st_params *p, int kind, uint64_t seed, void *d_u1, void *d_u2,
                       void *stream);

/* FP64 residual norm ||weight*S(u) - u||_2 over the interior of the F32 or
 * F64 field d_u (JACOBI operator), reduced on the GPU in FP64.  Blocks until
 * the result is on the host. */
st_status stencil_residual_norm(const st_params *p, const void *d_u, double *norm,
                                void *stream);

/* Mixed precision (p->dtype BF16 or F16, JACOBI operator, 3D or 2D):
 * p->iterations sweeps of u = x + e where x is an FP32 base field holding the
 * Dirichlet boundary, e a 16-bit correction and d a 16-bit defect.  Every
 * p->restart sweeps the correction is folded into x in FP32 and the defect
 * d = G(x) - x is recomputed in FP32.  Each correction sweep is
 *     e' = round16(widen(d) + weight * sum of widen(e) neighbours)
 * d_x (FP32 in: the initial field; scratch afterwards), d_e1, d_e2, d_d
 * (16-bit, nx*ny*nz each, scratch) and d_u (FP32 out, the final x + e; also
 * scratch during the run).  Bit-exact with oracle/stencil_oracle.c. */
st_status stencil_dispatch_mixed(const st_params *p, void *d_x, void *d_e1, void *d_e2,
                                 void *d_d, void *d_u, void *stream, float *elapsed_ms);

/* ------------------------------------------------------------------ */
/* Batched sweeps of many independent grids (3D, FP32, radius 1-2).     */
/* ------------------------------------------------------------------ */

/* One grid of a batch; the per-problem contract is stencil_dispatch's:
 * `iterations` sweeps ping-ponging between u1 and u2, result in u1 for an
 * even count, boundary shell of both buffers read-only.  In ST_BATCH_GRAPH a
 * problem whose two buffers hold the same shell takes passes of two fused
 * sweeps, which leave an intermediate state (not the previous sweep) in the
 * other buffer's interior, as stencil_dispatch's streaming form does. */
typedef struct st_problem {
    int nx, ny, nz;
    int radius;
    int iterations;
    int op;               /* st_op */
    double weight;        /* JACOBI weight, 0 = 1/(6*radius) */
    void *u1, *u2;        /* device buffers, nx*ny*nz floats each */
} st_problem;

/* Autotuned tile shape for a class of problems: the first rule with the
 * problem's radius and max_cells >= nx*ny*nz applies. */
typedef struct st_tile_rule {
    int radius;
    long long max_cells;
    int coarsen_y;        /* rows per tile frame (0 = as many as fit): the
                             plane-chunk form, used when every problem has
                             radius 1..2, (1+2R)*nx <= 4096, under 2^29
                             cells and all share one op; otherwise rows per
                             wave of the march form (1, 2 or 4) */
    int zchunk;           /* planes per tile (0 = whole z extent)         */
} st_tile_rule;

typedef enum st_batch_mode {
    ST_BATCH_GRAPH = 0,       /* one hipGraph replaying a batched launch per pass  */
    ST_BATCH_PERSISTENT = 1,  /* one persistent cooperative launch, a grid barrier
                                 between passes, captured in a hipGraph and
                                 replayed                                          */
    ST_BATCH_SEQUENTIAL = 2   /* stencil_dispatch per problem (baseline)           */
} st_batch_mode;

/* Advance every problem of the batch by its iteration count.  `stream`
 * must not be the null stream for ST_BATCH_GRAPH (a private stream is used
 * when it is NULL).  elapsed_ms (optional) covers the launches on the
 * device: in the plane-chunk form the comparison of each problem's two
 * shells (equal shells let whole rows be stored, shell cells with the value
 * they hold, and let two sweeps fuse into one pass) and the sweeps.
 * ST_BATCH_GRAPH and ST_BATCH_PERSISTENT keep the instantiated graph of the
 * last batch per device and stream and replay it when the next dispatch
 * has the same problems (buffers, shapes, iteration counts), tile rules and
 * mode; on the dispatch that captures and instantiates it, elapsed_ms
 * includes that host work (the stream waits meanwhile).  No mode may be
 * called on a stream under capture (ST_EINVAL): the table goes up from
 * host memory and the shell comparison is read back before the launches.
 * On the null stream ST_BATCH_PERSISTENT launches directly, without a
 * graph.  After a
 * fused pass the buffer that does not hold the result has an intermediate
 * state in its interior (its shell is untouched). */
st_status stencil_dispatch_batched(const st_problem *problems, int nproblems,
                                   const st_tile_rule *rules, int nrules, int mode,
                                   void *stream, float *elapsed_ms);

/* Human-readable status and the last error detail recorded on this thread. */
const char *stencil_status_str(st_status s);
const char *stencil_last_error(void);

/* ------------------------------------------------------------------ */
/* Host utilities for callers without their own HIP runtime plumbing    */
/* (the CLI, tests, bench).  Thin wrappers over hipMalloc and friends.  */
/* ------------------------------------------------------------------ */
st_status st_device_count(int *count);
st_status st_device_name(int device, char *buf, int len);
st_status st_malloc(void **ptr, uint64_t bytes);
st_status st_free(void *ptr);
st_status st_memcpy_htod(void *dst, const void *src, uint64_t bytes, void *stream);
st_status st_memcpy_dtoh(void *dst, const void *src, uint64_t bytes, void *stream);
st_status st_memcpy_dtod(void *dst, const void *src, uint64_t bytes, void *stream);
st_status st_memset(void *dst, int value, uint64_t bytes, void *stream);
st_status st_stream_create(void **stream);
st_status st_stream_destroy(void *stream);
st_status st_stream_sync(void *stream);
st_status st_device_sync(void);
/* Streaming copy of n floats (n a multiple of 4, 16-byte aligned buffers):
 * the bandwidth probe that calibrates HBM counters against the sweeps. */
st_status st_copy_f32(void *dst, const void *src, uint64_t n, void *stream);
/* The 16-bit storage conversions the mixed form uses (dtype ST_BF16 or
 * ST_F16), elementwise over n values: round to nearest even, and the exact
 * widening back.  No reference counterpart (the reference has no 16-bit
 * storage); for preparing and reading mixed-precision fields. */
st_status stencil_round16(const float *d_in, uint16_t *d_out, uint64_t n, int dtype, void *stream);
st_status stencil_widen16(const uint16_t *d_in, float *d_out, uint64_t n, int dtype, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* THESIS_AMD_STENCIL_ABI_H */
