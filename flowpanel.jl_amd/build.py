"""Build the gfx950 shared library libflowpanel_hip.so in-tree with hipcc.

Every HIP source under csrc/ is compiled for gfx950 only and linked into one
shared library exporting the C ABI of include/flowpanel_hip.h.  Packed FP32
instructions are disabled (``-packed-fp32-ops``); hipcc's host pass warns that
it does not know that target feature, which is harmless.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflowpanel_hip.so")
SOURCES = ["fp_influence.hip", "fp_dense.hip", "fp_adjoint.hip", "fp_post.hip", "fp_abi.hip", "fp_lu_nopiv.hip", "fp_assemble_edges.hip", "fp_field_trails.hip", "fp_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build(verbose=False, jobs=None):
    """Compile all sources (in parallel) and link the shared library."""
    with tempfile.TemporaryDirectory() as tmp:
        objs, procs = [], []
        for src in SOURCES:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            objs.append(obj)
            cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        failed = []
        for cmd, p in procs:
            out, _ = p.communicate()
            out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
            if p.returncode != 0:
                failed.append("%s\n%s" % (" ".join(cmd), out))
            elif verbose and out.strip():
                print(out)
        if failed:
            raise RuntimeError("HIP compilation failed:\n" + "\n".join(failed))
        tmp_lib = LIB + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp_lib] + objs, verbose)
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
