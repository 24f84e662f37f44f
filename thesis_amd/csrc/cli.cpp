// `stencil` command-line harness: the process contract of the reference's
// main() (src/main.cu:10-103) over the C ABI.
//
//   * knobs are flags named after the reference's -D macros
//     (Makefile:21-71): --dim --dimensions --radius --iterations --block-x
//     --block-y --block-z --coarsen-x --smem --padded --register --coop
//     --heuristic --smem-pad --debug, plus --nx/--ny/--nz, --op, --dtype,
//     --variant, --init, --seed, --coarsen-y, --temporal-steps;
//   * invalid knobs exit with status 1 (check_early_exit,
//     include/stencil_utils.h:113-124);
//   * the sweep loop is timed with HIP events and the time in milliseconds
//     is the last stdout line, printed "%.4f\n" (src/main.cu:89), which is
//     what the tuner parses;
//   * --debug compares the final field with the solution file
//     solutions/solution_<DIM>_<IT>_<R>_<D>d (include/stencil_utils.h:16-76):
//     raw values (float32, or float64 with --dtype f64), x fastest, no
//     header.  That name is the reference's for its own problem (THESIS
//     operator, FP32, the reference initial condition, a cubic grid); a run
//     that changes any of these appends what differs (the extents, _jacobi,
//     _f64, _seeded<seed>), so a cached solution is only ever compared with
//     the problem that made it.  A file of the wrong length is refused.
//     When the file is missing it is produced first by the BASE kernel
//     family on the GPU (the reference runs its CPU loop stencil_cpu() here,
//     src/main.cu:18-23); a CPU-made solution, such as the reference's own
//     solution files, is checked when it is placed under that name.  The
//     check prints "rms error = %f" and fails when the sum of squared
//     differences exceeds 1e-5 (include/stencil_error_checker.h:6-20).
#include <sys/stat.h>
#include <sys/types.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stencil_abi.h"

static void usage()
{
    std::fprintf(stderr,
                 "usage: stencil [--dim N | --nx N --ny N --nz N] [--dimensions 2|3] [--radius R]\n"
                 "               [--iterations I] [--block-x X --block-y Y --block-z Z] [--heuristic 0|1]\n"
                 "               [--coarsen-x C] [--coarsen-y C] [--smem] [--padded] [--register] [--coop]\n"
                 "               [--variant base|smem|smem_padded|smem_register|coop|auto]\n"
                 "               [--op thesis|jacobi] [--dtype f32|f64] [--init reference|seeded]\n"
                 "               [--seed S] [--temporal-steps T] [--form F] [--launch-per-pass]\n"
                 "               [--smem-pad P] [--debug] [--save]\n");
}

static bool file_exists(const std::string &f)
{
    struct stat b;
    return stat(f.c_str(), &b) == 0;
}

static int die(st_status s, const char *what)
{
    std::fprintf(stderr, "%s: %s (%s)\n", what, stencil_status_str(s), stencil_last_error());
    return 1;
}

int main(int argc, char **argv)
{
    st_params p;
    stencil_params_default(&p);
    bool smem = false, padded = false, reg = false, coop = false, debug = false, save = false;
    int variant = -1, init = -1;
    unsigned long long seed = 0x5EED0002ULL;
    int dim = -1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](void) -> const char * {
            if (i + 1 >= argc) { usage(); std::exit(1); }
            return argv[++i];
        };
        if (a == "--dim") dim = std::atoi(next());
        else if (a == "--nx") p.nx = std::atoi(next());
        else if (a == "--ny") p.ny = std::atoi(next());
        else if (a == "--nz") p.nz = std::atoi(next());
        else if (a == "--dimensions") p.dimensions = std::atoi(next());
        else if (a == "--radius") p.radius = std::atoi(next());
        else if (a == "--iterations") p.iterations = std::atoi(next());
        else if (a == "--block-x") p.block_x = std::atoi(next());
        else if (a == "--block-y") p.block_y = std::atoi(next());
        else if (a == "--block-z") p.block_z = std::atoi(next());
        else if (a == "--coarsen-x") p.coarsen_x = std::atoi(next());
        else if (a == "--coarsen-y") p.coarsen_y = std::atoi(next());
        else if (a == "--smem-pad") p.smem_pad = std::atoi(next());
        else if (a == "--heuristic") p.heuristic = std::atoi(next());
        else if (a == "--temporal-steps") p.temporal_steps = std::atoi(next());
        else if (a == "--launch-per-pass") p.launch_per_pass = 1;
        else if (a == "--form") p.temporal_form = std::atoi(next());
        else if (a == "--smem") smem = true;
        else if (a == "--padded") padded = true;
        else if (a == "--register") reg = true;
        else if (a == "--coop") coop = true;
        else if (a == "--debug") debug = true;
        else if (a == "--save") save = true;
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 0);
        else if (a == "--variant") {
            std::string v = next();
            if (v == "base") variant = ST_V_BASE;
            else if (v == "smem") variant = ST_V_SMEM;
            else if (v == "smem_padded") variant = ST_V_SMEM_PADDED;
            else if (v == "smem_register") variant = ST_V_SMEM_REGISTER;
            else if (v == "coop") variant = ST_V_COOP;
            else if (v == "auto") variant = ST_V_AUTO;
            else { usage(); return 1; }
        } else if (a == "--op") {
            std::string v = next();
            if (v == "jacobi") p.op = ST_OP_JACOBI;
            else if (v == "thesis") p.op = ST_OP_THESIS;
            else { usage(); return 1; }
        } else if (a == "--dtype") {
            std::string v = next();
            if (v == "f64") p.dtype = ST_F64;
            else if (v == "f32") p.dtype = ST_F32;
            else { usage(); return 1; }
        } else if (a == "--init") {
            std::string v = next();
            if (v == "seeded") init = ST_INIT_SEEDED;
            else if (v == "reference") init = ST_INIT_REFERENCE;
            else { usage(); return 1; }
        } else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            usage();
            return 1;
        }
    }
    if (dim > 0) p.nx = p.ny = p.nz = dim;
    if (p.dimensions == 2) p.nz = 1;
    // get_kernel() selection (src/stencil_dispatch.cu:14-42)
    if (variant >= 0) p.variant = variant;
    else if (coop) p.variant = ST_V_COOP;
    else if (smem && padded) p.variant = ST_V_SMEM_PADDED;
    else if (smem && reg) p.variant = ST_V_SMEM_REGISTER;
    else if (smem) p.variant = ST_V_SMEM;
    else p.variant = ST_V_BASE;
    if (init < 0) init = p.op == ST_OP_JACOBI ? ST_INIT_SEEDED : ST_INIT_REFERENCE;

    st_status s = stencil_validate(&p);
    if (s != ST_OK) return die(s, "invalid configuration");
    st_launch_info info;
    stencil_launch_config(&p, &info);
    static const char *names[] = {"base", "smem", "smem_padded", "smem_register", "coop", "auto"};
    std::printf("Version: %s (DIMENSIONS=%d NX=%d NY=%d NZ=%d BLOCK_X=%d BLOCK_Y=%d BLOCK_Z=%d "
                "RADIUS=%d COARSEN_X=%d)\n",
                names[info.variant], p.dimensions, p.nx, p.ny, p.dimensions == 3 ? p.nz : 1,
                info.block_x, info.block_y, info.block_z, p.radius, info.coarsen_x);

    int ndev = 0;
    if ((s = st_device_count(&ndev)) != ST_OK) return die(s, "device");
    // one stream of our own for all the work (the reference uses the default
    // stream; cooperative launches are not accepted on the null stream
    // between event records)
    void *strm = nullptr;
    if ((s = st_stream_create(&strm)) != ST_OK) return die(s, "stream");
    const size_t n = (size_t)p.nx * p.ny * (p.dimensions == 3 ? p.nz : 1);
    const size_t es = p.dtype == ST_F64 ? 8 : 4;
    void *d1 = nullptr, *d2 = nullptr;
    if ((s = st_malloc(&d1, n * es)) != ST_OK) return die(s, "alloc");
    if ((s = st_malloc(&d2, n * es)) != ST_OK) return die(s, "alloc");
    if ((s = stencil_init(&p, init, seed, d1, d2, strm)) != ST_OK) return die(s, "init");
    int in_u2 = 0;
    float ms = 0.f;
    if ((s = stencil_dispatch(&p, d1, d2, strm, &in_u2, &ms)) != ST_OK) return die(s, "dispatch");
    std::vector<unsigned char> host(n * es);
    if ((s = st_memcpy_dtoh(host.data(), in_u2 ? d2 : d1, n * es, strm)) != ST_OK) return die(s, "copy");
    if ((s = st_stream_sync(strm)) != ST_OK) return die(s, "sync");

    // SUFFIX of include/stencil_utils.h:14, then whatever makes this problem
    // differ from the reference's own
    std::string suffix = "_" + std::to_string(p.nx) + "_" + std::to_string(p.iterations) + "_" +
                         std::to_string(p.radius) + "_" + std::to_string(p.dimensions) + "d";
    const bool cubic = p.ny == p.nx && (p.dimensions == 2 || p.nz == p.nx);
    if (!cubic) {
        suffix += "_" + std::to_string(p.nx) + "x" + std::to_string(p.ny);
        if (p.dimensions == 3) suffix += "x" + std::to_string(p.nz);
    }
    if (p.op == ST_OP_JACOBI) suffix += "_jacobi";
    if (p.dtype == ST_F64) suffix += "_f64";
    if (init == ST_INIT_SEEDED) {
        char sb[32];
        std::snprintf(sb, sizeof sb, "_seeded%llx", seed);
        suffix += sb;
    }
    const int nz_print = p.dimensions == 3 ? p.nz : 1;
    bool error = false;
    if (debug) {
        const std::string sol = "solutions/solution" + suffix;
        if (!file_exists(sol)) {
            st_params q = p;
            q.variant = ST_V_BASE;
            q.heuristic = 1;
            void *r1 = nullptr, *r2 = nullptr;
            if ((s = st_malloc(&r1, n * es)) != ST_OK || (s = st_malloc(&r2, n * es)) != ST_OK)
                return die(s, "alloc");
            if ((s = stencil_init(&q, init, seed, r1, r2, strm)) != ST_OK) return die(s, "init");
            int r_in2 = 0;
            if ((s = stencil_dispatch(&q, r1, r2, strm, &r_in2, nullptr)) != ST_OK)
                return die(s, "solution");
            std::vector<unsigned char> ref(n * es);
            if ((s = st_memcpy_dtoh(ref.data(), r_in2 ? r2 : r1, n * es, strm)) != ST_OK) return die(s, "copy");
            if ((s = st_stream_sync(strm)) != ST_OK) return die(s, "sync");
            st_free(r1);
            st_free(r2);
            mkdir("solutions", 0700);
            FILE *f = std::fopen(sol.c_str(), "wb");
            if (!f) { std::perror(sol.c_str()); return 1; }
            const bool ok = std::fwrite(ref.data(), 1, ref.size(), f) == ref.size();
            if (std::fclose(f) != 0 || !ok) { std::perror(sol.c_str()); return 1; }
            std::printf("saving %s: nx = %d ny = %d nz = %d iterations: %d\n", sol.c_str(), p.nx, p.ny,
                        nz_print, p.iterations);
        }
        std::vector<unsigned char> ref(n * es);
        struct stat sb;
        FILE *f = nullptr;
        const bool size_ok = stat(sol.c_str(), &sb) == 0 && (size_t)sb.st_size == ref.size();
        if (size_ok) f = std::fopen(sol.c_str(), "rb");
        if (!f || std::fread(ref.data(), 1, ref.size(), f) != ref.size()) {
            if (f) std::fclose(f);
            std::printf("Unable to load file: %s\n", sol.c_str());
            return 1;
        }
        std::fclose(f);
        std::printf("reading %s: nx = %d ny = %d nz = %d\n", sol.c_str(), p.nx, p.ny, nz_print);
        double err = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (es == 4) {
                // the FP32 difference squared in FP32, accumulated in FP64
                // (include/stencil_error_checker.h:15)
                float a, b;
                std::memcpy(&a, &host[i * 4], 4);
                std::memcpy(&b, &ref[i * 4], 4);
                const float d = a - b;
                err += (double)(d * d);
            } else {
                double a, b;
                std::memcpy(&a, &host[i * 8], 8);
                std::memcpy(&b, &ref[i * 8], 8);
                const double d = a - b;
                err += d * d;
            }
        }
        std::printf("rms error = %f \n", std::sqrt(err / (double)n));
        error = !(err <= 10e-6);
    }
    if (save) {
        mkdir("results", 0700);
        const std::string res = "results/result" + suffix;
        FILE *f = std::fopen(res.c_str(), "wb");
        if (!f) { std::perror(res.c_str()); return 1; }
        const bool ok = std::fwrite(host.data(), 1, host.size(), f) == host.size();
        if (std::fclose(f) != 0 || !ok) { std::perror(res.c_str()); return 1; }
        std::printf("saving %s: nx = %d ny = %d nz = %d iterations: %d\n", res.c_str(), p.nx, p.ny, nz_print,
                    p.iterations);
    }
    std::printf("%.4f\n", ms);
    st_free(d1);
    st_free(d2);
    st_stream_destroy(strm);
    return error ? 1 : 0;
}
