"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/cp_oracle.cpp into oracle/liboracle.so with
``g++ -O2 -ffp-contract=off -fopenmp``.  -ffp-contract=off keeps every
``a*b + c`` as two roundings, matching the reference's Rust arithmetic.

The reference itself is Rust and cannot be built here (no cargo/rustc, and its
scirs2-core / scirs2-linalg 0.3.0 dependencies are not vendored), so there is
no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cp_oracle.cpp")
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-shared", "-fPIC",
           SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
