"""Build recipe for the CPU restatement library (test infrastructure only).

    python oracle/build.py

produces ``oracle/libamoracle.so``.  The flags keep IEEE operations exactly as
written (``-ffp-contract=off``), with explicit ``fma()`` where the source asks
for one, so the GPU engine's FP64 path can reproduce it bit for bit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "am_oracle.cpp")
OUT = os.path.join(HERE, "libamoracle.so")


def build(verbose=False):
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-mfma",
           "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
           SRC, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
