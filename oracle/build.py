"""Build recipe of the oracle's native part: the C++/OpenMP CPU port
(oracle/cpu_port.cpp -> oracle/lib/libtenet_cpu_port.so).

TEST INFRASTRUCTURE ONLY.  Targets x86-64-v3 (AVX2 + FMA) so the library
built here also runs on any current x86 host.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libtenet_cpu_port.so")


def build(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["g++", "-O3", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=fast", "-fopenmp", "-fPIC", "-shared",
           "-Wall", "-Wextra", os.path.join(HERE, "cpu_port.cpp"), "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
