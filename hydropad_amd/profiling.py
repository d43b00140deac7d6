"""Small driver for profiler runs (rocprofv3) of the sweep kernels.

    python -m hydropad_amd.profiling --size 512 --steps 2 --precision fp64 [--copy-mib 1024]

Uploads the planar Sod shock tube, runs `steps` full steps (x, y, z sweeps,
speed and timedet on the device) and, with --copy-mib, one plain
device-to-device hipMemcpy of that many MiB, whose counter readings
calibrate FETCH_SIZE / WRITE_SIZE against bytes actually moved (run with
HSA_ENABLE_SDMA=0 so the copy is a blit kernel the profiler can count).
"""
import argparse
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed", "fp64_exact"))
    ap.add_argument("--ic", default="sod", choices=("sod", "random"))
    ap.add_argument("--copy-mib", type=int, default=0)
    args = ap.parse_args()

    import hydropad_amd as hp
    from hydropad_amd import ic
    N = args.size
    st = ic.sod(N) if args.ic == "sod" else ic.random_smooth(N, 7)
    print(f"profiling: {args.ic} {N}^3, {args.steps} steps, {args.precision}", flush=True)
    with hp.Box(N, args.precision) as box:
        box.upload(st)
        dt, t = box.timedet(box.speed(), 0.8)
        t0 = time.perf_counter()
        box.evolve(1, args.steps, 0.8, t, dt)
        print(f"profiling: {args.steps} steps in {time.perf_counter() - t0:.3f} s", flush=True)
    if args.copy_mib > 0:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = ctypes.c_size_t(args.copy_mib * (1 << 20))
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(a), nbytes) == 0 and hip.hipMalloc(ctypes.byref(b), nbytes) == 0
        assert hip.hipMemset(a, 1, nbytes) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0          # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(a)
        hip.hipFree(b)
        print(f"profiling: copied {args.copy_mib} MiB device-to-device "
              f"({nbytes.value} bytes read, {nbytes.value} written)", flush=True)
    del st
    np.zeros(1)


if __name__ == "__main__":
    main()
