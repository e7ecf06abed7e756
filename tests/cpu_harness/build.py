"""Build tests/cpu_harness/libbvhcpu.so: the library's BVH logic compiled for
the host (test infrastructure)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbvhcpu.so")


def build():
    src = os.path.join(HERE, "bvh_cpu.hip")
    cmd = ["/opt/rocm/bin/hipcc", "--cuda-host-only", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
           "-o", LIB + ".tmp", src]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
