"""Build recipe for libso3ft.so (HIP, gfx950) -- run ``python -m pysofft_amd.build``.

The library is built in-tree next to this file so that it travels with the
source tree; nothing is installed.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libso3ft.so")
ARCH = os.environ.get("SO3FT_ARCH", "gfx950")

SOURCES = ["so3ft_api.hip", "dwt.hip", "fft.hip"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # never emit packed FP32 ALU ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    extra = [s for s in sorted(os.listdir(CSRC)) if s.endswith(".hip") and s not in SOURCES]
    srcs += [os.path.join(CSRC, s) for s in extra]
    tmp = LIB + ".tmp"
    cmd = ["/opt/rocm/bin/hipcc"] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
