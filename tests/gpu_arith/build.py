"""Build tests/gpu_arith/libarith_check.so (the device op check) in tree,
with the product's own HIP flags."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libarith_check.so")


def build():
    from mdz_amd import build as B
    hipcc = os.path.join(B.ROCM, "bin", "hipcc")
    cmd = [hipcc, *B.HIP_FLAGS, "-shared", os.path.join(HERE, "k_arith_check.hip"), "-o", LIB]
    B._run(cmd)
    return LIB


if __name__ == "__main__":
    print(build())
