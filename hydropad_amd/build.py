"""Build libhydropad_amd.so in-tree with hipcc for gfx950.

    python -m hydropad_amd.build

Compiles hp_kernels.hip + hp_capi.hip into one shared library next to this
file.  FP contraction is off so the FP64 sweep keeps the reference's
operation order; packed FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhydropad_amd.so")
ARCH = os.environ.get("HP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hp_kernels.hip", "hp_capi.hip"]
HEADERS = ["hp_core.h", "hp_kernels.h", os.path.join("..", "..", "include", "hydropad_gpu.h")]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result"]


def build(verbose=False):
    """Compile the HIP library (always recompiles)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + flags() + ["-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libhydropad_amd.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
