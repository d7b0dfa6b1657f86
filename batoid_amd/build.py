"""Build the gfx950 HIP library ``libbatoid_amd.so`` in-tree.

Each ``csrc/*.hip`` file is compiled for gfx950 with hipcc and the objects
are linked into one shared library next to this file.  The trace kernels are
compiled without floating-point contraction (``-ffp-contract=off``) so that
every FP64 result is bit-identical to the reference's non-FMA CPU build.

Usage: ``python -m batoid_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbatoid_amd.so")
SOURCES = ["bt_trace.hip", "bt_field.hip", "bt_huygens.hip", "bt_mlp.hip", "bt_raygen.hip",
           "bt_zernike.hip", "bt_split.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_resources=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        flags = list(COMMON_FLAGS)
        if verbose_resources:
            flags += ["-Rpass-analysis=kernel-resource-usage"]
        _run([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
