"""Build the C programs of tests/capi against include/mdzg.h and the in-tree
libmdzg.so (gcc, C99): abi_layout (no device calls), render_mdz, and
shim_pool (mdz_amd/shim/fractal_gpu.c, the line-callback shim MDZ would add,
built against the test-side declarations of tests/capi/mdz_side and driven
by a row-claiming thread pool)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROGS = ("abi_layout", "render_mdz")


def build():
    lib = os.path.join(ROOT, "mdz_amd")
    for p in PROGS:
        subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                        "-o", os.path.join(HERE, p), os.path.join(HERE, p + ".c"),
                        "-L", lib, "-lmdzg", "-Wl,-rpath," + lib], check=True)
    mp = os.environ.get("MDZG_MP_PREFIX", "/opt/conda")
    syslib = os.environ.get("MDZG_SYS_LIBDIR", "/usr/lib/x86_64-linux-gnu")
    subprocess.run(["gcc", "-std=gnu99", "-O2", "-Wall", "-Wextra", "-pthread",
                    "-I", os.path.join(HERE, "mdz_side"), "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(mp, "include"), "-o", os.path.join(HERE, "shim_pool"),
                    os.path.join(ROOT, "mdz_amd", "shim", "fractal_gpu.c"),
                    os.path.join(HERE, "shim_pool.c"), "-L", lib, "-lmdzg",
                    "-Wl,-rpath," + lib, f"-L{syslib}", "-l:libmpfr.so.6", "-l:libgmp.so.10"],
                   check=True)


if __name__ == "__main__":
    build()
