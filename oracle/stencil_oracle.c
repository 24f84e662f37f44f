/*
 * CPU oracle for the iterative star-stencil sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * It restates the reference CPU path of mrfarstad/thesis:
 *   - initialize_host_region{,_2d,_3d}   include/stencil_initializer.h:6-52
 *   - stencil()                          src/stencils.cu:8-19
 *   - accumulate_prev / accumulate_next  src/stencil_accumulators.cu:6-16
 *   - accumulate_global_{i,j,k}_*        src/stencil_accumulators.cu:46-74
 *   - cpu_stencil{,_2d,_3d}              src/stencil_cpu_kernel.cu:4-25
 *   - stencil_cpu() ping-pong driver     src/stencil_cpu.cu:20-43
 *   - check_domain_errors{,_2d,_3d}      include/stencil_error_checker.h:6-54
 * and adds the operator/initialisation forms the reference lacks:
 *   - JACOBI operator  u' = w * (sum of the 2*D*R axis neighbours)
 *   - counter-based seeded boundary initialisation
 *   - FP64 sweep and FP64 residual norm
 *   - the mixed-precision (16-bit storage, FP32 arithmetic) defect-correction
 *     scheme, bit-for-bit as the GPU implements it.
 *
 * Bit-exactness rules that the HIP kernels follow as well:
 *   - neighbour sum order: i-prev (d=R..1), i-next (d=1..R), j-prev, j-next,
 *     k-prev, k-next, accumulated left to right into a value that starts at 0;
 *   - THESIS: out = sum / (float)(2*D*R) - centre, IEEE division;
 *   - JACOBI: out = sum * w (one IEEE multiply);
 *   - no FMA contraction (built with -ffp-contract=off), denormals kept.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   or_init_reference_*    closed-form count of 1.0 cells; reference binary.
 *   or_sweep_* THESIS      golden outputs of the compiled reference stencil_cpu
 *                          (oracle/_ref, fixtures in tests/golden/), hand values
 *                          exact in binary, independent numpy brute force.
 *   or_sweep_* JACOBI      constant / linear / discrete-harmonic fixed points,
 *                          maximum principle, convergence to an FP64 DST-I
 *                          direct solve.
 *   or_residual_norm_*     zero for a discrete-harmonic field, numpy formula on
 *                          random fields.
 *   or_sse_f32             closed form on hand-made fields.
 *   or_mixed_*             FP64 error bounds and an independent numpy
 *                          restatement on small grids.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OP_THESIS 0
#define OP_JACOBI 1

/* ------------------------------------------------------------------ */
/* seeded counter-based generator                                      */
/* ------------------------------------------------------------------ */

/* splitmix64 finaliser (Steele, Lea, Flood 2014; public-domain constants). */
static inline uint64_t or_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* U(seed, idx): 24 random bits scaled by 2^-24, exactly representable in
 * FP32, uniform on [0,1). */
float or_uniform(uint64_t seed, uint64_t idx)
{
    uint64_t h = or_splitmix64(seed ^ (idx * 0x9E3779B97F4A7C15ULL));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

static inline int or_in_shell(long i, long j, long k, long nx, long ny, long nz,
                              int dims, int r)
{
    if (i < r || i >= nx - r) return 1;
    if (j < r || j >= ny - r) return 1;
    if (dims == 3 && (k < r || k >= nz - r)) return 1;
    return 0;
}

/* Reference initialisation: outermost layer 1.0, everything else 0.0.
 * include/stencil_initializer.h:6-52 (1.0 only at index 0 and N-1 of each
 * axis, independent of RADIUS). */
void or_init_reference_f32(float *u, long nx, long ny, long nz, int dims)
{
    if (dims == 2) nz = 1;
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                long idx = i + j * nx + k * nx * ny;
                int edge = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
                if (dims == 3) edge = edge || k == 0 || k == nz - 1;
                u[idx] = edge ? 1.0f : 0.0f;
            }
}

/* Seeded initialisation: every cell of the R-thick boundary shell gets
 * U(seed, idx), the interior is 0. */
void or_init_seeded_f32(float *u, long nx, long ny, long nz, int dims, int r,
                        uint64_t seed)
{
    if (dims == 2) nz = 1;
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                long idx = i + j * nx + k * nx * ny;
                u[idx] = or_in_shell(i, j, k, nx, ny, nz, dims, r)
                             ? or_uniform(seed, (uint64_t)idx) : 0.0f;
            }
}

void or_init_seeded_f64(double *u, long nx, long ny, long nz, int dims, int r,
                        uint64_t seed)
{
    if (dims == 2) nz = 1;
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                long idx = i + j * nx + k * nx * ny;
                u[idx] = or_in_shell(i, j, k, nx, ny, nz, dims, r)
                             ? (double)or_uniform(seed, (uint64_t)idx) : 0.0;
            }
}

/* ------------------------------------------------------------------ */
/* one sweep                                                            */
/* ------------------------------------------------------------------ */

/* One cell, FP32, in the reference's neighbour order
 * (src/stencils.cu:8-19, src/stencil_accumulators.cu:6-16,46-74).
 * The x stride of the j axis is nx (the reference's j_next uses NY, which is
 * the same value for the cubic grids it supports). */
static inline float or_cell_f32(const float *in, long idx, long sy, long sz,
                                int dims, int r, int op, float w, float coeff)
{
    float u = 0.0f;
    for (int d = r; d >= 1; d--) u += in[idx - d];
    for (int d = 1; d <= r; d++) u += in[idx + d];
    for (int d = r; d >= 1; d--) u += in[idx - d * sy];
    for (int d = 1; d <= r; d++) u += in[idx + d * sy];
    if (dims == 3) {
        for (int d = r; d >= 1; d--) u += in[idx - d * sz];
        for (int d = 1; d <= r; d++) u += in[idx + d * sz];
    }
    if (op == OP_THESIS) {
        float q = u / coeff;
        return q - in[idx];
    }
    return u * w;
}

static inline double or_cell_f64(const double *in, long idx, long sy, long sz,
                                 int dims, int r, int op, double w, double coeff)
{
    double u = 0.0;
    for (int d = r; d >= 1; d--) u += in[idx - d];
    for (int d = 1; d <= r; d++) u += in[idx + d];
    for (int d = r; d >= 1; d--) u += in[idx - d * sy];
    for (int d = 1; d <= r; d++) u += in[idx + d * sy];
    if (dims == 3) {
        for (int d = r; d >= 1; d--) u += in[idx - d * sz];
        for (int d = 1; d <= r; d++) u += in[idx + d * sz];
    }
    if (op == OP_THESIS) {
        double q = u / coeff;
        return q - in[idx];
    }
    return u * w;
}

static void or_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* The 3D interior loop of one sweep with the cell function CELL at radius
 * RR (k, j, i order; the OpenMP split is over k). */
#define OR_SWEEP3(CELL, RR)                                                     \
    do {                                                                        \
        _Pragma("omp parallel for schedule(static)")                            \
        for (long k = (RR); k < nz - (RR); k++)                                 \
            for (long j = (RR); j < ny - (RR); j++)                             \
                for (long i = (RR); i < nx - (RR); i++) {                       \
                    long idx = i + j * sy + k * sz;                             \
                    out[idx] = CELL(in, idx, sy, sz, 3, (RR), op, w, coeff);    \
                }                                                               \
    } while (0)

/* cpu_stencil_3d / cpu_stencil_2d (src/stencil_cpu_kernel.cu:4-19): interior
 * [R, N-R) on every axis, loop order k, j, i.  Boundary cells of `out` are
 * not written.  The OpenMP split over k (or j) does not change any cell's
 * operation sequence, so the result is bitwise independent of nthreads. */
void or_sweep_f32(const float *in, float *out, long nx, long ny, long nz,
                  int dims, int r, int op, float w, int nthreads)
{
    const long sy = nx, sz = nx * ny;
    const float coeff = (float)(2 * dims * r);
    or_set_threads(nthreads);
    if (dims == 3) {
        /* the radius as a constant for the common radii lets the compiler
         * unroll the neighbour loops; the operation order is the same */
        switch (r) {
        case 1: OR_SWEEP3(or_cell_f32, 1); break;
        case 2: OR_SWEEP3(or_cell_f32, 2); break;
        case 4: OR_SWEEP3(or_cell_f32, 4); break;
        default: OR_SWEEP3(or_cell_f32, r); break;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (long j = r; j < ny - r; j++)
            for (long i = r; i < nx - r; i++) {
                long idx = i + j * sy;
                out[idx] = or_cell_f32(in, idx, sy, 0, 2, r, op, w, coeff);
            }
    }
}

void or_sweep_f64(const double *in, double *out, long nx, long ny, long nz,
                  int dims, int r, int op, double w, int nthreads)
{
    const long sy = nx, sz = nx * ny;
    const double coeff = (double)(2 * dims * r);
    or_set_threads(nthreads);
    if (dims == 3) {
        switch (r) {
        case 1: OR_SWEEP3(or_cell_f64, 1); break;
        case 2: OR_SWEEP3(or_cell_f64, 2); break;
        case 4: OR_SWEEP3(or_cell_f64, 4); break;
        default: OR_SWEEP3(or_cell_f64, r); break;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (long j = r; j < ny - r; j++)
            for (long i = r; i < nx - r; i++) {
                long idx = i + j * sy;
                out[idx] = or_cell_f64(in, idx, sy, 0, 2, r, op, w, coeff);
            }
    }
}

/* stencil_cpu() loop (src/stencil_cpu.cu:32-35): `iterations` sweeps with a
 * pointer swap after each.  Returns 1 when the final field is in u2 (odd
 * iteration count), 0 when it is in u1. */
int or_run_f32(float *u1, float *u2, long nx, long ny, long nz, int dims,
               int r, int op, float w, int iterations, int nthreads)
{
    float *a = u1, *b = u2, *t;
    for (int it = 0; it < iterations; it++) {
        or_sweep_f32(a, b, nx, ny, nz, dims, r, op, w, nthreads);
        t = a; a = b; b = t;
    }
    return a == u2;
}

int or_run_f64(double *u1, double *u2, long nx, long ny, long nz, int dims,
               int r, int op, double w, int iterations, int nthreads)
{
    double *a = u1, *b = u2, *t;
    for (int it = 0; it < iterations; it++) {
        or_sweep_f64(a, b, nx, ny, nz, dims, r, op, w, nthreads);
        t = a; a = b; b = t;
    }
    return a == u2;
}

/* ------------------------------------------------------------------ */
/* residual norm (FP64 accumulation)                                    */
/* ------------------------------------------------------------------ */

/* ||r||_2 over interior cells, r = w*S(u) - u with S the neighbour sum
 * evaluated in FP64 from the stored values.  Accumulation in FP64 in
 * k, j, i order per thread block of k; the parallel reduction order makes
 * the last bits thread-count dependent, which the tests tolerate. */
double or_residual_norm_f32(const float *u, long nx, long ny, long nz,
                            int dims, int r, double w, int nthreads)
{
    const long sy = nx, sz = nx * ny;
    double acc = 0.0;
    or_set_threads(nthreads);
    if (dims == 2) nz = 1;
    long k0 = dims == 3 ? r : 0, k1 = dims == 3 ? nz - r : 1;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long k = k0; k < k1; k++)
        for (long j = r; j < ny - r; j++)
            for (long i = r; i < nx - r; i++) {
                long idx = i + j * sy + k * sz;
                double s = 0.0;
                for (int d = r; d >= 1; d--) s += (double)u[idx - d];
                for (int d = 1; d <= r; d++) s += (double)u[idx + d];
                for (int d = r; d >= 1; d--) s += (double)u[idx - d * sy];
                for (int d = 1; d <= r; d++) s += (double)u[idx + d * sy];
                if (dims == 3) {
                    for (int d = r; d >= 1; d--) s += (double)u[idx - d * sz];
                    for (int d = 1; d <= r; d++) s += (double)u[idx + d * sz];
                }
                double res = w * s - (double)u[idx];
                acc += res * res;
            }
    return sqrt(acc);
}

double or_residual_norm_f64(const double *u, long nx, long ny, long nz,
                            int dims, int r, double w, int nthreads)
{
    const long sy = nx, sz = nx * ny;
    double acc = 0.0;
    or_set_threads(nthreads);
    if (dims == 2) nz = 1;
    long k0 = dims == 3 ? r : 0, k1 = dims == 3 ? nz - r : 1;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long k = k0; k < k1; k++)
        for (long j = r; j < ny - r; j++)
            for (long i = r; i < nx - r; i++) {
                long idx = i + j * sy + k * sz;
                double s = 0.0;
                for (int d = r; d >= 1; d--) s += u[idx - d];
                for (int d = 1; d <= r; d++) s += u[idx + d];
                for (int d = r; d >= 1; d--) s += u[idx - d * sy];
                for (int d = 1; d <= r; d++) s += u[idx + d * sy];
                if (dims == 3) {
                    for (int d = r; d >= 1; d--) s += u[idx - d * sz];
                    for (int d = 1; d <= r; d++) s += u[idx + d * sz];
                }
                double res = w * s - u[idx];
                acc += res * res;
            }
    return sqrt(acc);
}

/* ------------------------------------------------------------------ */
/* check_domain_errors (include/stencil_error_checker.h:6-54)           */
/* ------------------------------------------------------------------ */

/* Sum of squared differences over the whole grid, FP64 accumulation in
 * index order (include/stencil_error_checker.h:6-21).  The reference prints
 * sqrt(sse/SIZE) as "rms error" and fails when sse > 10e-6. */
double or_sse_f32(const float *a, const float *b, long n)
{
    double err = 0.0;
    for (long i = 0; i < n; i++) {
        /* the reference squares the FP32 difference in FP32,
         * (h_u1[idx]-h_u2[idx])*(h_u1[idx]-h_u2[idx]), and adds the product
         * to its FP64 sum */
        float d = a[i] - b[i];
        err += (double)(d * d);
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* 16-bit storage with FP32 arithmetic and defect correction            */
/* ------------------------------------------------------------------ */

/* Round-to-nearest-even FP32 -> BF16 (NaN kept quiet) and back. */
static inline uint16_t or_f32_to_bf16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40);
    uint32_t lsb = (x >> 16) & 1u;
    x += 0x7fffu + lsb;
    return (uint16_t)(x >> 16);
}

static inline float or_bf16_to_f32(uint16_t h)
{
    uint32_t x = (uint32_t)h << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* IEEE binary16 conversions, round to nearest even, subnormals kept. */
static inline uint16_t or_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* overflow -> inf */
    if (ax < 0x38800000u) {                                    /* subnormal or 0 */
        if (ax < 0x33000000u) return (uint16_t)sign;             /* < 2^-25 -> 0 */
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;                               /* 14 + (113 - e) */
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t e = (ax >> 23) - 112u;
    uint32_t m = ax & 0x7fffffu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

static inline float or_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {
            float f = (float)m * (1.0f / 16777216.0f); /* m * 2^-24, exact */
            memcpy(&x, &f, 4);
            x |= sign;
        }
    } else if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &x, 4);
    return f;
}

uint16_t or_round16(float f, int is_f16) { return is_f16 ? or_f32_to_f16(f) : or_f32_to_bf16(f); }
float or_widen16(uint16_t h, int is_f16) { return is_f16 ? or_f16_to_f32(h) : or_bf16_to_f32(h); }

void or_round16_array(const float *in, uint16_t *out, long n, int is_f16)
{
    for (long i = 0; i < n; i++) out[i] = or_round16(in[i], is_f16);
}

void or_widen16_array(const uint16_t *in, float *out, long n, int is_f16)
{
    for (long i = 0; i < n; i++) out[i] = or_widen16(in[i], is_f16);
}

/* Mixed-precision JACOBI iteration, written u = x + e:
 *   x  FP32 base field (boundary shell holds the Dirichlet data),
 *   e  16-bit correction (boundary shell 0),
 *   d  16-bit defect  d = G(x) - x  on the interior (0 on the shell),
 * with G the JACOBI sweep.  Since G is affine, G(x + e) = G(x) + G0(e) with
 * G0 the homogeneous sweep, so one sweep of u is one correction sweep
 *     e' = d + G0(e)                 (FP32 arithmetic, stored in 16 bits)
 * Every `restart` sweeps the correction is folded back in FP32:
 *     x <- x + e,  e <- 0,  d <- G(x) - x   (defect in FP32, stored 16-bit)
 * Cell update order (bit-exact with the GPU):
 *     s = neighbour sum of widen(e) in reference order
 *     e'= round16( widen(d) + s * w )
 * Defect pass (FP32): t = G(x) cell as or_cell_f32 (JACOBI); d = round16(t - x).
 * Fold pass: x = x + widen(e).
 * The final field u = x + widen(e) (FP32) is written to u_out.
 * Returns the number of fold passes performed. */
int or_mixed_run(float *x, uint16_t *e1, uint16_t *e2, uint16_t *d, float *u_out,
                 long nx, long ny, long nz, int dims, int r, float w,
                 int iterations, int restart, int is_f16, int nthreads)
{
    const long sy = nx, sz = nx * ny;
    if (dims == 2) nz = 1;
    const long n = nx * ny * nz;
    const uint16_t zero16 = or_round16(0.0f, is_f16);
    long k0 = dims == 3 ? r : 0, k1 = dims == 3 ? nz - r : 1;
    int folds = 0;
    or_set_threads(nthreads);

    /* initial state: e = 0, d = G(x) - x */
    for (long i = 0; i < n; i++) { e1[i] = zero16; e2[i] = zero16; d[i] = zero16; }
#pragma omp parallel for schedule(static)
    for (long k = k0; k < k1; k++)
        for (long j = r; j < ny - r; j++)
            for (long i = r; i < nx - r; i++) {
                long idx = i + j * sy + k * sz;
                float t = or_cell_f32(x, idx, sy, sz, dims, r, OP_JACOBI, w, 0.0f);
                d[idx] = or_round16(t - x[idx], is_f16);
            }

    uint16_t *ea = e1, *eb = e2, *et;
    for (int it = 1; it <= iterations; it++) {
#pragma omp parallel for schedule(static)
        for (long k = k0; k < k1; k++)
            for (long j = r; j < ny - r; j++)
                for (long i = r; i < nx - r; i++) {
                    long idx = i + j * sy + k * sz;
                    float s = 0.0f;
                    for (int q = r; q >= 1; q--) s += or_widen16(ea[idx - q], is_f16);
                    for (int q = 1; q <= r; q++) s += or_widen16(ea[idx + q], is_f16);
                    for (int q = r; q >= 1; q--) s += or_widen16(ea[idx - q * sy], is_f16);
                    for (int q = 1; q <= r; q++) s += or_widen16(ea[idx + q * sy], is_f16);
                    if (dims == 3) {
                        for (int q = r; q >= 1; q--) s += or_widen16(ea[idx - q * sz], is_f16);
                        for (int q = 1; q <= r; q++) s += or_widen16(ea[idx + q * sz], is_f16);
                    }
                    float sw = s * w;
                    float v = or_widen16(d[idx], is_f16) + sw;
                    eb[idx] = or_round16(v, is_f16);
                }
        et = ea; ea = eb; eb = et;
        if (it % restart == 0 && it < iterations) {
            /* fold: x += e, e = 0 (both buffers), then new defect */
            for (long i = 0; i < n; i++) {
                x[i] = x[i] + or_widen16(ea[i], is_f16);
                ea[i] = zero16;
                eb[i] = zero16;
            }
#pragma omp parallel for schedule(static)
            for (long k = k0; k < k1; k++)
                for (long j = r; j < ny - r; j++)
                    for (long i = r; i < nx - r; i++) {
                        long idx = i + j * sy + k * sz;
                        float t = or_cell_f32(x, idx, sy, sz, dims, r, OP_JACOBI, w, 0.0f);
                        d[idx] = or_round16(t - x[idx], is_f16);
                    }
            folds++;
        }
    }
    for (long i = 0; i < n; i++) u_out[i] = x[i] + or_widen16(ea[i], is_f16);
    return folds;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
