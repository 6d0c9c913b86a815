"""Build the b12x HIP library in-tree for gfx950.

Two flavours share the sources:
  libb12x.so       parity build: FP64 contraction off, IEEE division (bit-exact with B12's host path)
  libb12x_fast.so  fast build: FP64 contraction (FMA) allowed in the map integrators
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("B12X_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-fno-fast-math", "-Wall", "-Wno-unused-function",
    # packed FP32 instructions stay disabled in every HIP build of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
FLAVOURS = {
    "libb12x.so": ["-ffp-contract=off"],
    "libb12x_fast.so": ["-ffp-contract=fast", "-DB12X_FAST_BUILD=1"],
}
SOURCES = [os.path.join(HERE, "csrc", "b12x_capi.hip"), os.path.join(HERE, "csrc", "b12x_forest.hip"),
           os.path.join(HERE, "csrc", "b12x_io.cpp")]


def build(flavours=None, verbose=False):
    for name in flavours or FLAVOURS:
        out = os.path.join(HERE, name)
        cmd = [HIPCC] + COMMON + FLAVOURS[name] + ["-o", out + ".tmp"] + SOURCES
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(sys.argv[1:] or None, verbose=True)
