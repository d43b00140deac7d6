"""Batched ensemble and serving loop over the C-ABI (include/hydropad_gpu.h,
hp_batch_* and hp_serve_run).

    b = Batch(max_boxes=64, hbm_budget=64 << 30)
    i = b.add(state, precision="mixed", nsteps=200)        # or add_sod(n, ...)
    b.use_graph(True); b.autotune(2); b.step(100); b.poll()
    out = b.download(i)

`serve(requests, ...)` runs the continuous-batching request loop and returns
per-request results (arrival / admission / completion times) and the
latency statistics.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import HpCosmo, check

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_I64, _U64, _SZ = ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t
PRECISIONS = {"fp64": 0, "mixed": 1}
FIELDS = ("p", "rho", "vx", "vy", "vz")


class HpBatchTuning(ctypes.Structure):
    _fields_ = [("x_chunks", _I), ("yz_segments", _I), ("waves_per_cu", _I), ("reserved", _I)]

    def as_dict(self):
        return {"x_chunks": self.x_chunks, "yz_segments": self.yz_segments, "waves_per_cu": self.waves_per_cu}


class HpBoxStatus(ctypes.Structure):
    _fields_ = [("steps_done", _I64), ("active", ctypes.c_int32), ("done", ctypes.c_int32), ("t", _D), ("dt", _D),
                ("newton_iters_max", _I64), ("newton_nonconv", _I64), ("neg_state_clamps", _I64)]


class HpRequest(ctypes.Structure):
    _fields_ = [("id", _I64), ("n", ctypes.c_int32), ("nsteps", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("newton_limit", ctypes.c_int32), ("seed", _U64), ("arrival_s", _D), ("cour", _D)]


class HpRequestResult(ctypes.Structure):
    _fields_ = [("id", _I64), ("arrival_s", _D), ("admit_s", _D), ("done_s", _D), ("steps", _I64), ("t", _D),
                ("newton_nonconv", _I64), ("rc", ctypes.c_int32), ("pad", ctypes.c_int32)]


class HpServeStats(ctypes.Structure):
    _fields_ = [("wall_s", _D), ("cell_updates", _D), ("p50_s", _D), ("p99_s", _D), ("mean_s", _D),
                ("batched_steps", _I64), ("max_concurrent", _I64), ("completed", _I64), ("peak_bytes", _D)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_bound = False


def _lib():
    global _bound
    L = _capi.lib()
    if not _bound:
        PB = _P
        L.hp_batch_box_bytes.argtypes = [_I, _I, _I, _I]
        L.hp_batch_box_bytes.restype = _SZ
        L.hp_batch_create.argtypes = [_I, _SZ, ctypes.POINTER(PB)]
        L.hp_batch_destroy.argtypes = [PB]
        L.hp_batch_destroy.restype = None
        L.hp_batch_add.argtypes = [PB, _I, _I, _I, _I, _I, _D, _D, _D, ctypes.POINTER(HpCosmo), _I] + [_P] * 5 + \
            [ctypes.POINTER(_I)]
        L.hp_batch_add_sod.argtypes = [PB, _I, _I, _U64, _D, _I, ctypes.POINTER(_I)]
        L.hp_sod_of_seed.argtypes = [_U64, ctypes.POINTER(_I), ctypes.POINTER(_D), ctypes.POINTER(_I)]
        L.hp_batch_remove.argtypes = [PB, _I]
        L.hp_batch_set_newton_limit.argtypes = [PB, _I, _I]
        L.hp_batch_download.argtypes = [PB, _I] + [_P] * 5
        L.hp_batch_step.argtypes = [PB, _I]
        L.hp_batch_sync.argtypes = [PB]
        L.hp_batch_poll.argtypes = [PB, ctypes.POINTER(HpBoxStatus), ctypes.POINTER(_I)]
        L.hp_batch_use_graph.argtypes = [PB, _I]
        L.hp_batch_set_tuning.argtypes = [PB, ctypes.POINTER(HpBatchTuning)]
        L.hp_batch_get_tuning.argtypes = [PB, ctypes.POINTER(HpBatchTuning)]
        L.hp_batch_bytes_used.argtypes = [PB]
        L.hp_batch_bytes_used.restype = _SZ
        L.hp_batch_autotune.argtypes = [PB, _I, ctypes.POINTER(HpBatchTuning), ctypes.POINTER(_D),
                                        ctypes.POINTER(_I)]
        L.hp_serve_run.argtypes = [ctypes.POINTER(HpRequest), _I, _SZ, _I, _I, ctypes.POINTER(_P),
                                   ctypes.POINTER(HpRequestResult), ctypes.POINTER(HpServeStats)]
        _bound = True
    return L


def box_bytes(shape, comoving=False):
    nx, ny, nz = shape
    return int(_lib().hp_batch_box_bytes(nx, ny, nz, 1 if comoving else 0))


def sod_of_seed(seed):
    """(axis, x0, flip) of the seeded Sod tube of Batch.add_sod / serve."""
    a, x0, f = _I(), _D(), _I()
    check("hp_sod_of_seed", _lib().hp_sod_of_seed(int(seed), ctypes.byref(a), ctypes.byref(x0), ctypes.byref(f)))
    return a.value, x0.value, f.value


class Batch:
    def __init__(self, max_boxes=64, hbm_budget=64 << 30):
        self.L = _lib()
        h = _P()
        check("hp_batch_create", self.L.hp_batch_create(int(max_boxes), int(hbm_budget), ctypes.byref(h)))
        self._h = h
        self.max_boxes = int(max_boxes)
        self.shapes = {}

    def close(self):
        if self._h:
            self.L.hp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, state, nsteps, precision="fp64", cour=0.8, cosmo=None, tstop=1e30, gamma=5.0 / 3.0, nbound=3):
        f = [np.asfortranarray(state[k], dtype=np.float64) for k in FIELDS]
        nx, ny, nz = f[0].shape
        i = _I()
        cp = ctypes.byref(cosmo) if cosmo is not None else None
        check("hp_batch_add", self.L.hp_batch_add(self._h, nx, ny, nz, nbound, PRECISIONS[precision], gamma, cour,
                                                  tstop, cp, int(nsteps), *[a.ctypes.data_as(_P) for a in f],
                                                  ctypes.byref(i)))
        self.shapes[i.value] = (nx, ny, nz)
        return i.value

    def add_sod(self, n, seed, nsteps, precision="fp64", cour=0.8):
        i = _I()
        check("hp_batch_add_sod", self.L.hp_batch_add_sod(self._h, int(n), PRECISIONS[precision], int(seed), cour,
                                                          int(nsteps), ctypes.byref(i)))
        self.shapes[i.value] = (n + 6,) * 3
        return i.value

    def remove(self, i):
        check("hp_batch_remove", self.L.hp_batch_remove(self._h, int(i)))
        self.shapes.pop(i, None)

    def set_newton_limit(self, i, nmax=0):
        check("hp_batch_set_newton_limit", self.L.hp_batch_set_newton_limit(self._h, int(i), int(nmax)))

    def download(self, i):
        out = {k: np.zeros(self.shapes[i], dtype=np.float64, order="F") for k in FIELDS}
        check("hp_batch_download", self.L.hp_batch_download(self._h, int(i),
                                                            *[out[k].ctypes.data_as(_P) for k in FIELDS]))
        return out

    def step(self, nsteps=1):
        check("hp_batch_step", self.L.hp_batch_step(self._h, int(nsteps)))

    def sync(self):
        check("hp_batch_sync", self.L.hp_batch_sync(self._h))

    def poll(self):
        arr = (HpBoxStatus * self.max_boxes)()
        n = _I()
        check("hp_batch_poll", self.L.hp_batch_poll(self._h, arr, ctypes.byref(n)))
        return {i: arr[i] for i in range(self.max_boxes) if arr[i].active}

    def run_until_done(self, chunk=1):
        while True:
            st = self.poll()
            if all(s.done for s in st.values()):
                return st
            self.step(chunk)

    def use_graph(self, on=True):
        check("hp_batch_use_graph", self.L.hp_batch_use_graph(self._h, 1 if on else 0))

    def set_tuning(self, x_chunks=0, yz_segments=0, waves_per_cu=0):
        t = HpBatchTuning(x_chunks, yz_segments, waves_per_cu, 0)
        check("hp_batch_set_tuning", self.L.hp_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self):
        t = HpBatchTuning()
        check("hp_batch_get_tuning", self.L.hp_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, steps_per_trial=2):
        t, ms, n = HpBatchTuning(), _D(), _I()
        check("hp_batch_autotune", self.L.hp_batch_autotune(self._h, int(steps_per_trial), ctypes.byref(t),
                                                            ctypes.byref(ms), ctypes.byref(n)))
        return t.as_dict(), ms.value, n.value

    def bytes_used(self):
        return int(self.L.hp_batch_bytes_used(self._h))


def serve(requests, hbm_budget=64 << 30, max_boxes=64, use_graph=True, keep_fields=()):
    """Run the request loop.  requests: list of dicts with n, nsteps,
    precision ("fp64"/"mixed"), seed, arrival_s (and optional id, cour,
    newton_limit).
    keep_fields: indices of requests whose final states are returned.
    Returns (results list of dicts, stats dict, {index: state})."""
    L = _lib()
    nreq = len(requests)
    arr = (HpRequest * max(nreq, 1))()
    for i, r in enumerate(requests):
        arr[i] = HpRequest(int(r.get("id", i)), int(r["n"]), int(r["nsteps"]), PRECISIONS[r["precision"]],
                           int(r.get("newton_limit", 0)), int(r["seed"]), float(r["arrival_s"]),
                           float(r.get("cour", 0.8)))
    res = (HpRequestResult * max(nreq, 1))()
    stats = HpServeStats()
    fields = {}
    ptrs = None
    if keep_fields:
        ptrs = (_P * (5 * nreq))()
        for i in keep_fields:
            n = requests[i]["n"] + 6
            fields[i] = {k: np.zeros((n, n, n), dtype=np.float64, order="F") for k in FIELDS}
            for j, k in enumerate(FIELDS):
                ptrs[5 * i + j] = fields[i][k].ctypes.data
    check("hp_serve_run", L.hp_serve_run(arr, nreq, int(hbm_budget), int(max_boxes), 1 if use_graph else 0, ptrs,
                                         res, ctypes.byref(stats)))
    out = [{n: getattr(res[i], n) for n, _ in HpRequestResult._fields_ if n != "pad"} for i in range(nreq)]
    return out, stats.as_dict(), fields
