"""Builds the C part of the oracle (liboracle_knn.so) with gcc.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "knn_oracle.c")
LIB = os.path.join(HERE, "liboracle_knn.so")


def build(force=True, quiet=False):
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-o", tmp, SRC, "-lm"]
    if not quiet:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
