"""Build recipe for the CPU oracle: oracle/fp_oracle.c -> oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY (the checker, never the product path).  The reference
(FLOWPanel.jl) is Julia; no Julia toolchain exists in this image, so there is
no oracle/_ref build: the oracle is this C restatement, pinned by the tests.
Importing this module builds the library.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fp_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
           "-o", OUT + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


build()
