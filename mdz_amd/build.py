"""Build mdz_amd/libmdzg.so in-tree: gfx950 kernels (hipcc) + C host runtime (gcc).

    python -m mdz_amd.build [-j JOBS]

The kernels are compiled once per word count N (k_rn.hip with -DMDZG_N=N), plus
the FP64 and GMP-path objects, all in parallel, then linked with the host
runtime against MPFR/GMP.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdzg.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
MP_PREFIX = os.environ.get("MDZG_MP_PREFIX", "/opt/conda")
NWORDS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-pragma-unroll-threshold=1000000",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = [
    "-O2", "-std=gnu99", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-ffp-contract=off", "-pthread",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{MP_PREFIX}/include",
]
# MPFR, GMP and zlib are linked from the system library directory, by
# soname, with no run path: a C program (MDZ itself) that links libmdzg.so
# then resolves the C++ runtime the ROCm libraries need from the system, not
# from a prefix that ships an older one.  The headers come from MP_PREFIX;
# the sonames (libmpfr.so.6, libgmp.so.10) fix the ABI.
SYS_LIB = os.environ.get("MDZG_SYS_LIBDIR", "/usr/lib/x86_64-linux-gnu")
LINK_LIBS = [f"-L{SYS_LIB}", "-l:libmpfr.so.6", "-l:libgmp.so.10", "-l:libz.so.1",
             "-lpthread", "-lm", "-ldl"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def jobs():
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    steps = []
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "k_gmp.hip"),
                   "-o", os.path.join(BUILD, "k_gmp.o")], "k_gmp.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "k_fp64.hip"),
                   "-o", os.path.join(BUILD, "k_fp64.o")], "k_fp64.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "k_colour.hip"),
                   "-o", os.path.join(BUILD, "k_colour.o")], "k_colour.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "noalloc.hip"),
                   "-o", os.path.join(BUILD, "noalloc.o")], "noalloc.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "k_probe.hip"),
                   "-o", os.path.join(BUILD, "k_probe.o")], "k_probe.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "graph.hip"),
                   "-o", os.path.join(BUILD, "graph.o")], "graph.o"))
    for n in sorted(NWORDS, reverse=True):   # largest (slowest) first
        steps.append(([hipcc, *HIP_FLAGS, f"-DMDZG_N={n}", "-c", os.path.join(CSRC, "k_rn.hip"),
                       "-o", os.path.join(BUILD, f"k_rn{n}.o")], f"k_rn{n}.o"))
    steps.append(([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "launch.hip"),
                   "-o", os.path.join(BUILD, "launch.o")], "launch.o"))
    for c in sorted(os.listdir(HOST)):
        if c.endswith(".c"):
            o = c[:-2] + ".o"
            steps.append((["gcc", *C_FLAGS, "-c", os.path.join(HOST, c),
                           "-o", os.path.join(BUILD, o)], o))
    return steps


def build(nproc=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    steps = jobs()
    nproc = nproc or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=nproc) as ex:
        futs = {ex.submit(_run, cmd): name for cmd, name in steps}
        for f in cf.as_completed(futs):
            out = f.result()
            if verbose and out.strip():
                print(futs[f], out, file=sys.stderr)
    objs = [os.path.join(BUILD, name) for _, name in steps]
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC",
          "-o", tmp, *objs, *LINK_LIBS])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
