/*
 * Harness that drives the reference CPU oracle (mrfarstad/thesis
 * src/stencil_cpu.cu:20-43) from its own sources.  TEST INFRASTRUCTURE ONLY:
 * oracle/ref_build.sh compiles it against the reference tree, with the
 * reference's compile-time knobs (-D DIM, DIMENSIONS, RADIUS, ITERATIONS),
 * and runs it inside oracle/_ref/, where stencil_cpu() writes its raw
 * float32 solution file solutions/solution_<DIM>_<IT>_<R>_<D>d
 * (include/stencil_utils.h:16-44).
 *
 * stencil_cpu() mallocs its second buffer without initialising it
 * (src/stencil_cpu.cu:26-27).  Forcing every allocation through mmap makes
 * that buffer zero-filled, which is what fresh memory holds in the
 * reference's own runs; the CPU oracle of this repository zeroes it
 * explicitly.
 */
#include <malloc.h>
#include "src/stencil_cpu.cu"

int main()
{
    mallopt(M_MMAP_THRESHOLD, 0);
    stencil_cpu();
    return 0;
}
