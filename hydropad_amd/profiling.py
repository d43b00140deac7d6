"""Small driver for profiler runs (rocprofv3) of the sweep kernels.

    python -m hydropad_amd.profiling --size 512 --steps 2 --precision fp64 [--copy-mib 1024]
    python -m hydropad_amd.profiling --summarize DB [DB ...] [--copy-mib 1024]
    python -m hydropad_amd.profiling --traffic FETCH_DB WRITE_DB --size 512 --precision fp64 --copy-mib 1024
    python -m hydropad_amd.profiling --valu SQ_DB --size 512 --precision fp64

Uploads the planar Sod shock tube, runs `steps` full steps (x, y, z sweeps,
speed and timedet on the device) and, with --copy-mib, one plain
device-to-device hipMemcpy of that many MiB, whose counter readings
calibrate FETCH_SIZE / WRITE_SIZE against bytes actually moved (run with
HSA_ENABLE_SDMA=0 so the copy is a blit kernel the profiler can count).
"""
import argparse
import collections
import sqlite3
import time

import numpy as np


def summarize(paths, copy_bytes=None, last=0):
    """Per-kernel summary of rocprofv3 result databases (kernel trace and/or
    counter collection).  Returns text; with copy_bytes, FETCH_SIZE and
    WRITE_SIZE are also given in bytes, scaled by what the same counter read
    for the calibration copy (the largest __amd_rocclr_copyBuffer dispatch).
    With last > 0, each sweep kernel's figures are averaged over its last
    `last` dispatches only (the late state of a long run)."""
    out = []
    for path in paths:
        db = sqlite3.connect(path)
        tabs = {r[0] for r in db.execute("select name from sqlite_master")}
        out.append(f"## {path}")
        if "kernels" in tabs and db.execute("select count(*) from kernels").fetchone()[0]:
            out.append("| kernel | calls | avg ms | total ms | VGPR | AGPR |")
            out.append("|---|---|---|---|---|---|")
            agg = collections.defaultdict(list)
            regs = {}
            for name, dur, vg, ag in db.execute("select name, duration, vgpr_count, accum_vgpr_count from kernels "
                                                "order by start"):
                agg[name].append(dur)
                regs[name] = (vg, ag)
            if last > 0:
                agg = {k: (v[-last:] if "sweep" in k else v) for k, v in agg.items()}
            for name, d in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
                out.append(f"| {name.split('(')[0][:80]} | {len(d)} | {np.mean(d) / 1e6:.4f} | "
                           f"{np.sum(d) / 1e6:.3f} | {regs[name][0]} | {regs[name][1]} |")
        if "counters_collection" in tabs and db.execute("select count(*) from counters_collection").fetchone()[0]:
            vals = collections.defaultdict(lambda: collections.defaultdict(list))
            for disp, name, cn, v in db.execute(
                    "select dispatch_id, kernel_name, counter_name, value from counters_collection "
                    "order by dispatch_id"):
                vals[name.split("(")[0]][cn].append(v)
            if last > 0:
                for k, d in vals.items():
                    if "sweep" in k:
                        for cn in d:
                            d[cn] = d[cn][-last:]
            scale = {}
            if copy_bytes:
                cp = [k for k in vals if "copyBuffer" in k]
                for cn in ("FETCH_SIZE", "WRITE_SIZE"):
                    if cp and cn in vals[cp[0]]:
                        scale[cn] = copy_bytes / (max(vals[cp[0]][cn]) * 1024.0)
            cnames = sorted({c for d in vals.values() for c in d})
            out.append("| kernel | dispatches | " + " | ".join(cnames) + " |")
            out.append("|---|---|" + "---|" * len(cnames))
            for name, d in vals.items():
                if not ("sweep" in name or "copyBuffer" in name or "speed" in name):
                    continue
                cells = []
                for cn in cnames:
                    if cn not in d:
                        cells.append("")
                        continue
                    v = float(np.max(d[cn])) if "copyBuffer" in name else float(np.mean(d[cn]))
                    txt = f"{v:.4g}"
                    if cn in scale:
                        txt += f" ({v * 1024.0 * scale[cn] / 1e9:.3f} GB)"
                    cells.append(txt)
                n = len(next(iter(d.values())))
                out.append(f"| {name[:80]} | {n} | " + " | ".join(cells) + " |")
            if scale:
                out.append("")
                out.append("calibration: " + ", ".join(f"{k} x{v:.3f} (bytes per KiB read)" for k, v in scale.items())
                           + f" from a {copy_bytes / 2**30:.0f} GiB device-to-device copy")
        out.append("")
    return "\n".join(out)


def traffic_entries(fetch_db, write_db, size, precision, copy_bytes):
    """HBM bytes per launch of each sweep kernel from one FETCH_SIZE and one
    WRITE_SIZE run, each counter scaled by what it read for the calibration
    copy (copy_bytes moved each way by its largest dispatch)."""
    def load(path, cn):
        db = sqlite3.connect(path)
        vals = collections.defaultdict(list)
        for name, v in db.execute("select kernel_name, value from counters_collection where counter_name=?", (cn,)):
            vals[name.split("(")[0]].append(v)
        return vals
    f, w = load(fetch_db, "FETCH_SIZE"), load(write_db, "WRITE_SIZE")
    cf = [k for k in f if "copyBuffer" in k][0]
    cw = [k for k in w if "copyBuffer" in k][0]
    fscale = copy_bytes / (max(f[cf]) * 1024.0)
    wscale = copy_bytes / (max(w[cw]) * 1024.0)
    out = {}
    alg = 80 * size ** 3
    for k in f:
        if "sweep" not in k or k not in w:
            continue
        short = "sweep_lane_x" if "lane_x" in k else ("sweep_stream<y>" if ", 2," in k else "sweep_stream<z>")
        fb = float(np.mean(f[k])) * 1024.0 * fscale
        wb = float(np.mean(w[k])) * 1024.0 * wscale
        out[f"{short}@{size}@{precision}"] = {
            "bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb), "write_bytes": round(wb),
            "alg_bytes": alg, "ratio_to_alg": round((fb + wb) / alg, 4),
            "fetch_scale_bytes_per_KiB": fscale, "write_scale_bytes_per_KiB": wscale, "kernel": k}
    return out


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD for 4
# cycles (16 lanes per cycle, the FP64 and unpacked-FP32 rate) at the 2.4 GHz
# peak engine clock
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def valu_entries(sq_db, size, precision):
    """VALU issue figures per launch of each sweep kernel from one SQ counter
    run: SQ_INSTS_VALU (wave instructions, averaged over every dispatch of
    the kernel) and the time the chip needs just to issue them,
    insts x 4 cycles / (1024 SIMDs x 2.4 GHz) -- the floor set by instruction
    count alone, whatever the memory traffic."""
    db = sqlite3.connect(sq_db)
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, cn, v in db.execute("select kernel_name, counter_name, value from counters_collection"):
        vals[name.split("(")[0]][cn].append(v)
    out = {}
    for k, d in vals.items():
        if "sweep" not in k or "SQ_INSTS_VALU" not in d:
            continue
        short = "sweep_lane_x" if "lane_x" in k else ("sweep_stream<y>" if ", 2," in k else "sweep_stream<z>")
        insts = float(np.mean(d["SQ_INSTS_VALU"]))
        out[f"{short}@{size}@{precision}"] = {
            "valu_insts_per_launch": round(insts), "dispatches": len(d["SQ_INSTS_VALU"]),
            "valu_insts_per_64_cells": round(insts / (size ** 3 / 64.0), 1),
            "issue_bound_ms": round(insts * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e3, 4), "kernel": k}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed", "fp64_exact"))
    ap.add_argument("--ic", default="sod", choices=("sod", "random", "cosmo"),
                    help="cosmo: the comoving workload (Gaussian random field, Mach 50, EdS from a = 1/51 "
                         "to 1/11 over 200 steps), run for --steps of its steps")
    ap.add_argument("--axis", type=int, default=0, choices=(0, 1, 2), help="shock-tube axis")
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--summarize", nargs="+", metavar="DB", help="print a summary of rocprofv3 databases and exit")
    ap.add_argument("--last", type=int, default=0,
                    help="with --summarize: average each sweep kernel over its last LAST dispatches")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="merge per-launch HBM bytes of the sweep kernels into --traffic-out and exit")
    ap.add_argument("--traffic-out", default="profiles/traffic.json")
    ap.add_argument("--valu", metavar="SQ_DB",
                    help="merge per-launch VALU issue figures of the sweep kernels into --valu-out and exit")
    ap.add_argument("--valu-out", default="profiles/valu.json")
    args = ap.parse_args()
    if args.traffic:
        import json
        import os
        tr = json.load(open(args.traffic_out)) if os.path.exists(args.traffic_out) else {}
        tr.update(traffic_entries(args.traffic[0], args.traffic[1], args.size, args.precision,
                                  args.copy_mib * (1 << 20)))
        json.dump(tr, open(args.traffic_out, "w"), indent=1)
        print(json.dumps({k: v["bytes_per_launch"] for k, v in tr.items()}, indent=1))
        return
    if args.valu:
        import json
        import os
        tr = json.load(open(args.valu_out)) if os.path.exists(args.valu_out) else {}
        tr.update(valu_entries(args.valu, args.size, args.precision))
        json.dump(tr, open(args.valu_out, "w"), indent=1)
        print(json.dumps({k: v["issue_bound_ms"] for k, v in tr.items()}, indent=1))
        return
    if args.summarize:
        print(summarize(args.summarize, args.copy_mib * (1 << 20) if args.copy_mib else None, args.last))
        return

    import hydropad_amd as hp
    from hydropad_amd import ic
    N = args.size
    if args.ic == "cosmo":
        st = ic.grf_box(N, 2026)
    else:
        st = ic.sod(N, axis=args.axis) if args.ic == "sod" else ic.random_smooth(N, 7)
    print(f"profiling: {args.ic} (axis {args.axis}) {N}^3, {args.steps} steps, {args.precision}", flush=True)
    with hp.Box(N, args.precision) as box:
        box.upload(st)
        t0 = time.perf_counter()
        if args.ic == "cosmo":
            ca, tstop = ic.eds_schedule(1.0 / 51.0, 1.0 / 11.0, 200)
            box.evolve_cosmo(1, args.steps, hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop))
        else:
            dt, t = box.timedet(box.speed(), 0.8)
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
