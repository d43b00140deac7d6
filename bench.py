#!/usr/bin/env python3
"""Benchmark of the directionally split PPM hydro step on one MI355X.

Workload (the metric's configuration): 512^3 periodic box, planar Sod shock
tube along x (rho,p = 1,1 for x<0.5, 0.125,0.1 otherwise, v = 0), gamma 5/3,
CFL 0.8, nrot sweep schedule, on-device speed + timedet every step.  The box
is advanced K steps in FP64 and K steps in the mixed FP32/FP64 form, each
from the same initial state; the metric is cell-updates/s (N^3 * K / wall
seconds) with the state resident in HBM, and the mixed form's per-field
relative L1/Linf against the FP64 run.

    python bench.py [--gpus 1] [--steps 300] [--warmup 5] [--size 512] [--no-extras]
                    [--legs cosmo,ensemble,serving,slab]

Prints one JSON line.  `roofline` is the dominant sweep kernel's achieved HBM
bandwidth (80 algorithmic bytes per cell per sweep / its average HIP-event
duration) against 8 TB/s; `cpu_baseline` times the oracle's C restatement of
the reference sweep (bit-identical to the reference's Fortran) on bounded
slabs of the same workload, on the host's cores and on one thread, with the
reference itself timed beside it where its build is present.  Beside the headline the line carries the other
workloads of the metric:
  cosmo_256   comoving 256^3 box, a = 1/51 -> 1/11 in 200 steps, FP64 (and mixed)
  ensemble    64 comoving boxes of sizes {32,48,64,96,128,192}^3, own seeds,
              P(k) slopes and rms, one batch, autotuned, hipGraph; FP64 and mixed
  serving     Poisson stream of Sod-tube requests (N log-uniform 32..256,
              20..200 steps, FP64 or mixed), continuous batching; p50/p99 latency
  zslab_share one GPU's z-slab of the 512^3 tube split over 8 GPUs (512x512x64,
              ghosts filled locally), FP64 and mixed
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_SWEEP = 80          # 5 FP64 fields read + 5 written
BYTES_PER_CELL_STEP = 3 * BYTES_PER_CELL_SWEEP
NB = 3
FIELDS = ("p", "rho", "vx", "vy", "vz")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_precision(hp, st0, N, precision, steps, warmup, cour):
    """Warm up, re-upload the initial state, time `steps` steps."""
    with hp.Box(N, precision) as box:
        box.upload(st0)
        vm = box.speed()
        dt0, t0 = box.timedet(vm, cour)
        if warmup > 0:
            box.evolve(1, warmup, cour, t0, dt0)
        box.upload(st0)
        box.params.dt, box.params.rdtath = dt0, dt0
        box.set_params(box.params)
        box.sync()
        box.sweep_timing_enable(True)
        box.reset_counters()
        t_start = time.perf_counter()
        t, dt = box.evolve(1, steps, cour, t0, dt0)      # synchronises on return
        wall = time.perf_counter() - t_start
        ms, launches = box.sweep_timing()
        cnt = box.counters()
        box.sweep_timing_enable(False)
        out = box.download()
    return {"wall": wall, "ms": ms, "launches": launches, "state": out, "t": t, "dt": dt,
            "newton_max": int(cnt.newton_iters_max), "nonconv": int(cnt.newton_nonconv)}


def errors(a, b):
    res = {}
    for k in FIELDS:
        x = a[k][NB:-NB, NB:-NB, NB:-NB]
        y = b[k][NB:-NB, NB:-NB, NB:-NB]
        den1 = np.abs(y).sum()
        den8 = np.abs(y).max()
        d = np.abs(x - y)
        res[k] = {"l1": float(d.sum() / den1) if den1 > 0 else float(d.sum()),
                  "linf": float(d.max() / den8) if den8 > 0 else float(d.max())}
    return res


def roofline_of(r, N, prec):
    names = ("sweep_lane_x", "sweep_stream<y>", "sweep_stream<z>")
    tot = [r["ms"][d] for d in range(3)]
    dom = int(np.argmax(tot))
    avg_ms = r["ms"][dom] / max(r["launches"][dom], 1)
    alg_bytes = BYTES_PER_CELL_SWEEP * N ** 3
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            tr = json.load(open(TRAFFIC_FILE))
            key = f"{names[dom]}@{N}@{prec}"
            if key in tr:
                traffic = tr[key]["bytes_per_launch"]
        except (ValueError, KeyError):
            traffic = None
    per_dir = {names[d]: round(r["ms"][d] / max(r["launches"][d], 1), 4) for d in range(3)}
    out = {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "alg_bytes_per_launch": alg_bytes, "avg_ms_per_launch": per_dir}
    # the issue floor: the dominant kernel's VALU instructions per launch (SQ
    # counters of the same workload, profiles/valu.json) at one wave64
    # instruction per 4 cycles on each of 1024 SIMDs at 2.4 GHz, against the
    # launch time measured here
    if os.path.exists(VALU_FILE):
        try:
            v = json.load(open(VALU_FILE)).get(f"{names[dom]}@{N}@{prec}")
            if v:
                out["valu_issue"] = {"insts_per_launch": v["valu_insts_per_launch"],
                                     "insts_per_64_cells": v["valu_insts_per_64_cells"],
                                     "issue_bound_ms": v["issue_bound_ms"],
                                     "frac": round(v["issue_bound_ms"] / avg_ms, 4)}
        except (ValueError, KeyError):
            pass
    return out


def cpu_baseline(N, budget_s=10.0):
    """CPU baseline on the host cores.

    The timed legs are the oracle's C restatement of the reference's STENCIL
    sweep ("port": same per-interface structure and arithmetic, bit-identical
    to the reference) on slabs of the workload (same dx = 1/N, the same
    shock-tube pencils), each run until ~budget_s: N x 64 x 64 on all the
    host's cores (OpenMP over pencils, at most 16 threads -- the CPU share of
    one GPU on the GPU host), which is `value`, and N x 32 x 32 on one thread.  The reference's own
    Fortran build (oracle/_ref) only runs cubic boxes -- its y/z Riemann
    gathers index rho3d(iaux,j,k) transposed (hydro/ppmsolver.F90:251,379) --
    so it is timed beside the port on one 64^3 step as a calibration."""
    try:
        from oracle import Oracle, Reference
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": str(e)}
    order = {1: (1, 2, 3), 2: (3, 2, 1), 3: (2, 3, 1), 4: (1, 3, 2), 5: (3, 1, 2), 6: (2, 1, 3)}

    def shock_tube(nx, ntr, ngrid):
        n = nx + 2 * NB
        st = {k: np.zeros((n, ntr + 2 * NB, ntr + 2 * NB), order="F") for k in FIELDS}
        x = (np.arange(n) - NB + 0.5) / nx
        st["rho"][...] = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
        st["p"][...] = np.where(x < 0.5, 1.0, 0.1)[:, None, None]
        P = O.params(ngrid=ngrid)
        O.timedet(P, O.speed(P, st), 0.8, 1e30, 0.0)
        return st, P

    O = Oracle()

    def timed(ntr, threads):
        O.set_threads(threads)
        st, P = shock_tube(N, ntr, N)
        t0 = time.perf_counter()
        nsteps = 0
        while time.perf_counter() - t0 < budget_s and nsteps < 100:
            nsteps += 1
            O.ppm_lxlylz(P, st, 1 if nsteps == 1 else nsteps % 6 + 1)
        el = time.perf_counter() - t0
        O.set_threads(1)
        return N * ntr * ntr * nsteps / el, nsteps, el

    # every host thread this process may use: the box grants one GPU's share of
    # the host and says so in OMP_NUM_THREADS; without it, every logical CPU
    ncpu = os.cpu_count() or 1
    try:
        threads = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu))))
    except ValueError:
        threads = ncpu
    v_all, n_all, el_all = timed(64, threads)
    v_one, n_one, el_one = timed(32, 1)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    out = {"value": round(v_all, 1), "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "host_cpu": model, "host_logical_cpus": ncpu,
           "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
           "sample": f"{N}x64x64 slab of the {N}^3 shock tube (dx=1/{N}), {n_all} steps "
                     f"(3 sweeps + periodic fill each) in {el_all:.1f} s on {threads} threads (OpenMP over pencils)",
           "single_thread": {"value": round(v_one, 1), "cores": 1,
                             "sample": f"{N}x32x32 slab, {n_one} steps in {el_one:.1f} s on one thread"}}
    if Reference.available():
        try:
            R = Reference()
            st_r, P_r = shock_tube(64, 64, 64)
            st_p = {k: v.copy(order="F") for k, v in st_r.items()}
            R.set_params(P_r)
            g0 = np.zeros((3,) + st_r["p"].shape, order="F")
            t1 = time.perf_counter()
            for d in order[1]:
                R.ppmsolver(st_r, d, gforce=g0)
                O.exchange_mesh(st_r)
            tr = time.perf_counter() - t1
            t1 = time.perf_counter()
            O.ppm_lxlylz(P_r, st_p, 1)
            tp = time.perf_counter() - t1
            same = all(np.array_equal(st_r[k], st_p[k]) for k in FIELDS)
            out["reference_calibration"] = {
                "reference_cell_updates_per_s": round(64 ** 3 / tr, 1),
                "port_cell_updates_per_s": round(64 ** 3 / tp, 1), "bit_identical": bool(same),
                "sample": "one step of the 64^3 shock tube, reference (flang -O2) vs port, one thread"}
        except Exception as e:  # the reference build is optional on the GPU host
            out["reference_calibration"] = {"error": str(e)}
    return out


def run_cosmo(hp, ic, N=256, steps=200):
    """Comoving box of the cosmological workload: cell-updates/s per precision,
    mixed-vs-FP64 error."""
    st = ic.grf_box(N, 2026)
    ca, tstop = ic.eds_schedule(1.0 / 51.0, 1.0 / 11.0, steps)
    out = {"workload": f"grf_{N}^3_eds_a1/51-1/11_{steps}steps_mach50", "box": N, "steps": steps}
    states = {}
    for prec in ("fp64", "mixed"):
        with hp.Box(N, prec) as box:
            s = hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop)
            box.upload(st)
            box.sync()
            t0 = time.perf_counter()
            box.evolve_cosmo(1, steps, s)
            wall = time.perf_counter() - t0
            states[prec] = box.download()
        out[f"{prec}_cell_updates_per_s"] = round(N ** 3 * steps / wall, 1)
        out[f"{prec}_ms_per_step"] = round(wall / steps * 1e3, 4)
        out["a_end"] = s.atnew
    out["mixed_vs_fp64_err"] = errors(states["mixed"], states["fp64"])
    return out


def ensemble_boxes(ic, nbox=64, seed=4242):
    rng = np.random.default_rng(seed)
    sizes = rng.choice([32, 48, 64, 96, 128, 192], size=nbox)
    boxes = []
    for i, n in enumerate(sizes):
        slope = float(rng.uniform(-2.5, -1.5))
        rms = float(rng.uniform(0.02, 0.1))
        boxes.append((int(n), 1000 + i, slope, rms))
    return boxes


def run_ensemble(hp, ic, steps=50, nbox=64):
    """64-box comoving ensemble in one batch (persistent per-direction kernels,
    hipGraph, autotuned tiling), FP64 and mixed, each from the same states.
    The tiling is tuned on a separate batch of the same boxes, so the timed
    steps are the ensemble's steps 1..steps whatever the tuner ran."""
    from hydropad_amd import batch
    spec = ensemble_boxes(ic, nbox)
    log(f"bench: ensemble ICs ({nbox} boxes)")
    states = [ic.grf_box(n, sd, slope=sl, rms=rms) for n, sd, sl, rms in spec]
    cells = sum(n ** 3 for n, _, _, _ in spec)
    ca, tstop = ic.eds_schedule(1.0 / 51.0, 1.0 / 11.0, 200)
    out = {"workload": f"{nbox}_comoving_boxes_sizes32-192_{steps}steps", "boxes": nbox,
           "cells": int(cells), "steps": steps, "size_histogram": {}}
    for n, _, _, _ in spec:
        out["size_histogram"][str(n)] = out["size_histogram"].get(str(n), 0) + 1
    finals = {}
    for prec in ("fp64", "mixed"):
        with batch.Batch(max_boxes=nbox, hbm_budget=200 << 30) as b:
            for st in states:
                b.add(st, nsteps=200, cosmo=hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop), precision=prec)
            best, tune_ms, nrun = b.autotune(steps_per_trial=1)
        with batch.Batch(max_boxes=nbox, hbm_budget=200 << 30) as b:
            ids = [b.add(st, nsteps=steps, cosmo=hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop),
                         precision=prec) for st in states]
            b.set_tuning(**best)
            b.use_graph(True)
            b.step(1)                                   # graph capture, untimed
            b.sync()
            lat = []
            t0 = time.perf_counter()
            for _ in range(steps - 1):
                t1 = time.perf_counter()
                b.step(1)
                b.sync()
                lat.append(time.perf_counter() - t1)
            wall = time.perf_counter() - t0
            st = b.poll()
            assert all(s.steps_done == steps for s in st.values())
            finals[prec] = [b.download(i) for i in ids]
        out[prec] = {"cell_updates_per_s": round(cells * (steps - 1) / wall, 1),
                     "ms_per_batched_step": round(wall / (steps - 1) * 1e3, 3),
                     "step_latency_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                     "step_latency_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                     "tuning": best, "autotune_steps": nrun}
    worst = {}
    for a, b_ in zip(finals["mixed"], finals["fp64"]):
        e = errors(a, b_)
        for k, v in e.items():
            w = worst.setdefault(k, {"l1": 0.0, "linf": 0.0})
            w["l1"] = max(w["l1"], v["l1"])
            w["linf"] = max(w["linf"], v["linf"])
    out["mixed_vs_fp64_worst_box_err"] = worst
    return out


def run_slab(hp, N=512, nz=64, steps=300, cour=0.8):
    """One GPU's share of the 512^3 tube split into z-slabs across 8 GPUs: a
    512 x 512 x 64 box (dx = 1/512, the same planar shock tube in x) stepped
    with its z ghosts filled locally, the work a rank would do between halo
    exchanges.  The exchange itself (3 planes x 5 fields each side before a
    z sweep, 2 x 3 x 5 x 8 x 512^2 B = 63 MB) is not run: one GPU here."""
    shape = (N + 2 * NB, N + 2 * NB, nz + 2 * NB)
    xc = (np.arange(shape[0]) - NB + 0.5) / N
    left = (xc < 0.5)[:, None, None]
    st = {"p": np.where(left, 1.0, 0.1), "rho": np.where(left, 1.0, 0.125)}
    for k in ("vx", "vy", "vz"):
        st[k] = np.zeros(shape)
    st = {k: np.asfortranarray(np.broadcast_to(v, shape), dtype=np.float64) for k, v in st.items()}
    out = {"workload": f"zslab_{N}x{N}x{nz}_sod_{steps}steps", "shape": [N, N, nz], "steps": steps,
           "halo_bytes_per_z_sweep": 2 * NB * 5 * 8 * N * N}
    for prec in ("fp64", "mixed"):
        with hp.Box(shape, prec) as box:
            box.upload(st)
            dt, t = box.timedet(box.speed(), cour)
            box.sync()
            t0 = time.perf_counter()
            box.evolve(1, steps, cour, t, dt)
            box.sync()
            wall = time.perf_counter() - t0
        out[f"{prec}_cell_updates_per_s"] = round(N * N * nz * steps / wall, 1)
        out[f"{prec}_ms_per_step"] = round(wall / steps * 1e3, 4)
    return out


def run_serving(nreq=48, rate=8.0, seed=5150):
    """Poisson stream of box evolutions through the continuous-batching loop."""
    from hydropad_amd import batch
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        n = int(round(np.exp(rng.uniform(np.log(32), np.log(256)))))
        reqs.append({"n": n, "nsteps": int(rng.integers(20, 201)), "precision": "mixed" if rng.random() < 0.5
                     else "fp64", "seed": int(rng.integers(1 << 62)), "arrival_s": t})
    res, stats, _ = batch.serve(reqs, hbm_budget=64 << 30, max_boxes=64, use_graph=True)
    lat = np.array([r["done_s"] - r["arrival_s"] for r in res if r["rc"] == 0])
    out = {"workload": f"poisson_{rate}req/s_{nreq}req_N32-256_steps20-200_sod_seeded",
           "requests": nreq, "rate_per_s": rate, "completed": int(stats["completed"]),
           "cell_updates_per_s": round(stats["cell_updates"] / stats["wall_s"], 1),
           "wall_s": round(stats["wall_s"], 3), "p50_latency_s": round(stats["p50_s"], 4),
           "p99_latency_s": round(stats["p99_s"], 4), "mean_latency_s": round(float(lat.mean()), 4),
           "batched_steps": int(stats["batched_steps"]), "max_concurrent": int(stats["max_concurrent"]),
           "peak_hbm_gb": round(stats["peak_bytes"] / 1e9, 3)}
    # The serving loop's own cost, apart from the arrival process: the same
    # requests all present at t = 0 through the serving loop (admission,
    # one-step-ahead queueing, completion) against one plain batch holding
    # them from the start and stepped until every box is done
    sat, st2, _ = batch.serve([dict(r, arrival_s=0.0) for r in reqs], hbm_budget=64 << 30, max_boxes=64,
                              use_graph=True)
    assert all(r["rc"] == 0 for r in sat)
    cells = sum(r["n"] ** 3 * r["nsteps"] for r in reqs)
    with batch.Batch(max_boxes=64, hbm_budget=64 << 30) as b:
        for r in reqs:
            b.add_sod(r["n"], r["seed"], r["nsteps"], r["precision"])
        b.use_graph(True)
        b.sync()
        t0 = time.perf_counter()
        b.run_until_done(chunk=8)
        b.sync()
        offline_s = time.perf_counter() - t0
    out["saturated"] = {"cell_updates_per_s": round(st2["cell_updates"] / st2["wall_s"], 1),
                        "wall_s": round(st2["wall_s"], 3),
                        "offline_batch_cell_updates_per_s": round(cells / offline_s, 1),
                        "offline_batch_wall_s": round(offline_s, 3),
                        "serving_vs_offline_batch": round(st2["cell_updates"] / st2["wall_s"] / (cells / offline_s), 4)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cour", type=float, default=0.8)
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--legs", default="cosmo,ensemble,serving,slab",
                    help="comma-separated other workloads to run beside the headline")
    args = ap.parse_args()
    legs = set() if args.no_extras else {x.strip() for x in args.legs.split(",") if x.strip()}
    known = {"cosmo", "ensemble", "serving", "slab"}
    if legs - known:
        ap.error(f"unknown legs: {sorted(legs - known)}")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a code path)"}))
        sys.exit(2)

    import hydropad_amd as hp
    from hydropad_amd import ic
    N, K = args.size, args.steps
    log(f"bench: {N}^3 shock tube, {K} steps, warmup {args.warmup}")
    st0 = ic.sod(N)
    results = {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        log(f"bench: running {prec}")
        results[prec] = run_precision(hp, st0, N, prec, K, args.warmup, args.cour)
        log(f"bench: {prec} wall {results[prec]['wall']:.3f} s")

    cups = {p: N ** 3 * K / r["wall"] for p, r in results.items()}
    err = None
    if "fp64" in results and "mixed" in results:
        err = errors(results["mixed"]["state"], results["fp64"]["state"])
    mixed_ok = err is not None and err["rho"]["l1"] <= 1e-5
    head = "mixed" if mixed_ok else ("fp64" if "fp64" in results else "mixed")
    r = results[head]
    line = {
        "metric": "cell-updates/s at 512^3, mixed vs FP64 at L1<=1e-5; ensemble p50/p99 latency",
        "value": round(cups[head], 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (Sod shock tube initial condition)",
        "config": {"workload": f"sod_{N}^3_cfl0.8_{K}steps", "box": N, "nbound": NB, "gamma": "5/3",
                   "path": head},
        "roofline": roofline_of(r, N, head),
        "step_roofline_frac": round(cups[head] * BYTES_PER_CELL_STEP / (HBM_PEAK_GBS * 1e9), 4),
        "fp64_cell_updates_per_s": round(cups["fp64"], 1) if "fp64" in cups else None,
        "mixed_cell_updates_per_s": round(cups["mixed"], 1) if "mixed" in cups else None,
        "fp64_roofline_frac": round(cups["fp64"] * BYTES_PER_CELL_STEP / (HBM_PEAK_GBS * 1e9), 4)
        if "fp64" in cups else None,
        "mixed_vs_fp64_err": err,
        "newton_iters_max": {p: v["newton_max"] for p, v in results.items()},
    }
    if "fp64" in results and head != "fp64":
        line["roofline_fp64"] = roofline_of(results["fp64"], N, "fp64")
    if "cosmo" in legs:
        log("bench: cosmological 256^3")
        line["cosmo_256"] = run_cosmo(hp, ic)
    if "ensemble" in legs:
        log("bench: ensemble")
        line["ensemble"] = run_ensemble(hp, ic)
    if "serving" in legs:
        log("bench: serving")
        line["serving"] = run_serving()
    if "slab" in legs:
        log("bench: z-slab share")
        line["zslab_share"] = run_slab(hp, N=N, steps=K, cour=args.cour)
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(N)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
