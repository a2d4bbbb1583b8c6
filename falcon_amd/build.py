"""Builds libfalcon_hip.so for gfx950 in-tree (falcon_amd/_lib/).

    python -m falcon_amd.build

Every HIP source is compiled with -ffp-contract=off (the FP64 path keeps the
reference's operation order, no fused multiply-add) and with packed-FP32
instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libfalcon_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["layout.cpp", "synthetic.cpp", "output.cpp", "kernels.hip", "tables.hip", "falcon.hip", "heat.hip", "cnn.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


# host-only sources: sin and cos stay separate libm calls (no sincos fusion), the
# same convention as the oracle, so setup tables are reproducible across compilers
HOST_ONLY = ["-fno-builtin-sin", "-fno-builtin-cos"]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT_DIR, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else HOST_ONLY
        cmd = [HIPCC] + FLAGS + lang + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
