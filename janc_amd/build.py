"""Build the gfx950 step library (libjanc_amd.so) in-tree with hipcc.

Regenerates the mechanism source from the packaged YAML first, so the
compiled reaction structure always matches ``mechanism.py``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjanc_amd.so")
ARCH = os.environ.get("JANC_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=" + ARCH,
    # packed FP32 arithmetic is kept off in every HIP build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False, extra=()):
    from .mechanism import Mechanism, write_generated
    write_generated(Mechanism(), CSRC)
    src = os.path.join(CSRC, "janc_step.hip")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + list(extra) + [src, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
