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

# (source, FP contraction): the exact sweep keeps the reference's operation
# order, so it is built without contraction; the throughput sweeps with it.
SOURCES = [("hp_kernels.hip", "off"), ("hp_kernels_exact.hip", "off"), ("hp_kernels_fast.hip", "fast"),
           ("hp_capi.hip", "off"), ("hp_batch_kernels.hip", "fast"), ("hp_batch.hip", "off"),
           ("hp_serve.hip", "off"), ("hp_noalloc.hip", "off"), ("hp_nbody.hip", "off")]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags(contract):
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=" + contract,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _compile(args):
    src, contract, obj, verbose = args
    cmd = [hipcc()] + flags(contract) + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    """Compile the HIP library (always recompiles; translation units in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = [(s, c, os.path.join(objdir, s.replace(".hip", ".o")), verbose) for s, c in SOURCES]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        results = list(ex.map(_compile, jobs))
    for src, rc, out in results:
        if rc != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc failed on {src}")
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [j[2] for j in jobs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libhydropad_amd.so failed")
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "hydropad_driver")


def flang():
    for c in ("/opt/rocm/llvm/bin/flang", shutil.which("flang") or ""):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module (fortran/hydropad_gpu.F90) and link the
    evolve_sys-shaped driver against libhydropad_amd.so.  Skipped, with a
    message, when no Fortran compiler is installed."""
    fc = flang()
    if fc is None:
        sys.stderr.write("build: no flang; skipping the Fortran module and driver\n")
        return None
    objdir = os.path.join(HERE, "_build")
    mod_o = os.path.join(objdir, "hydropad_gpu_f.o")
    steps = [[fc, "-O2", "-fPIC", "-module-dir", objdir, "-c", os.path.join(FORTRAN, "hydropad_gpu.F90"),
              "-o", mod_o],
             [fc, "-O2", "-module-dir", objdir, os.path.join(FORTRAN, "hydropad_driver.F90"), mod_o,
              "-L" + HERE, "-lhydropad_amd", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"]]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed: " + " ".join(cmd))
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
