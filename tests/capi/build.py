"""Build the C programs of tests/capi against include/mdzg.h and the in-tree
libmdzg.so (gcc, C99): abi_layout (no device calls) and render_mdz."""
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


if __name__ == "__main__":
    build()
