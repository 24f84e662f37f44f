"""Build the gfx950 shared library ``thesis_amd/libthesis_stencil.so``.

Every HIP source is compiled by hipcc for gfx950 only, with
``-ffp-contract=off`` (bit-exact operation order with the CPU oracle) and
packed-FP32 instruction selection disabled.  Objects go to
``thesis_amd/_build/``; the library is linked in-tree so it travels with the
source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libthesis_stencil.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
    "-ffp-contract=off", "-fno-fast-math",
    f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-pass-failed",
]
SOURCES = [
    "dispatch.cpp",
    "sweep_base.hip",
    "sweep_march.hip",
    "sweep_coop.hip",
    "sweep_tblock.hip", "sweep_stream.hip", "sweep_stream2d.hip", "sweep_rows.hip",
    "batched.hip",
    "unet.hip",
    "init_reduce.hip",
    "mixed.hip",
    "noalloc.hip",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + [f"--offload-arch={ARCH}", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the `stencil` CLI (process contract of the reference's main())
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    cli = os.path.join(HERE, "bin", "stencil")
    cmd = [HIPCC] + COMMON + ["-o", cli, os.path.join(CSRC, "cli.cpp"), "-L" + HERE,
                              "-lthesis_stencil", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {cli}")
    return LIB


if __name__ == "__main__":
    build()
