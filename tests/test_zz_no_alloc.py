"""hp_ppmsolver_noalloc -- one sweep on caller-owned device buffers, with the
scalars and counters in the library's per-GPU scratch array -- against
hp_ppmsolver on the same inputs: bit for bit on the interior and the face
ghosts, with equal counters, on calls in a row where each later call is on a
smaller box (the scratch keeps the previous call's contents, so a stale
scalar or counter would show)."""
import ctypes

import numpy as np
import pytest

import hydropad_amd as hp
from helpers import FIELDS, NB, copy_state
from hydropad_amd import _capi, ic
from hydropad_amd.hydro import PRECISIONS

pytestmark = pytest.mark.gpu


def box_sweep(st, P, d, precision, newton_limit=0):
    with hp.Box(st["p"].shape, precision) as box:
        box.set_params(P)
        box.set_newton_limit(newton_limit)
        box.upload(copy_state(st))
        rc = box.L.hp_ppmsolver(box._h, d)
        out = box.download()
        c = box.counters()
    return rc, out, c


def hip():
    """The HIP runtime the library itself links (device buffers for the test;
    the call under test allocates nothing)."""
    _capi.lib()
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            H = ctypes.CDLL(name)
            break
        except OSError:
            continue
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


H2D, D2H = 1, 2


def noalloc_sweep(st, P, d, precision, newton_limit=0):
    H = hip()
    shape = st["p"].shape
    nbytes = st["p"].size * 8
    bufs = []
    for _ in range(10):
        b = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(b), nbytes) == 0
        bufs.append(b)
    try:
        nan = np.full(st["p"].size, np.nan)
        for i, k in enumerate(FIELDS):
            a = np.ascontiguousarray(st[k].ravel(order="F"))
            assert H.hipMemcpy(bufs[i], a.ctypes.data, nbytes, H2D) == 0
            assert H.hipMemcpy(bufs[5 + i], nan.ctypes.data, nbytes, H2D) == 0
        c = hp.HpCounters()
        rc = _capi.lib().hp_ppmsolver_noalloc(ctypes.byref(P), *bufs, None, *shape, NB, d, PRECISIONS[precision],
                                              newton_limit, ctypes.byref(c), None)
        out = {}
        for i, k in enumerate(FIELDS):
            a = np.empty(st["p"].size)
            assert H.hipMemcpy(a.ctypes.data, bufs[5 + i], nbytes, D2H) == 0
            out[k] = a.reshape(shape, order="F")
    finally:
        for b in bufs:
            H.hipFree(b)
    return rc, out, c


def faces_and_interior(a):
    """The interior and the periodic face-ghost slabs the sweep writes."""
    nx, ny, nz = a.shape
    I = (slice(NB, nx - NB), slice(NB, ny - NB), slice(NB, nz - NB))
    parts = [a[I]]
    for ax in range(3):
        for g in (slice(0, NB), slice(a.shape[ax] - NB, None)):
            sl = list(I)
            sl[ax] = g
            parts.append(a[tuple(sl)])
    return parts


CALLS = [
    ("rnd20-x-fp64", lambda: ic.random_smooth(20, 7), 1, "fp64", 0),
    ("sod16-y-exact", lambda: ic.sod(16, axis=1), 2, "fp64_exact", 0),
    ("rnd-13x11x9-z-mixed", lambda: {k: np.asfortranarray(v[:19, :17, :15])
                                     for k, v in ic.periodic_fill(ic.random_smooth(16, 9)).items()}, 3, "mixed", 0),
    ("sod10-x-limit2", lambda: ic.sod(10), 1, "fp64", 2),
]


def test_noalloc_sweep_matches_ppmsolver():
    for name, make, d, precision, lim in CALLS:
        print(f"noalloc {name}", flush=True)
        st = make()
        ic.periodic_fill(st)
        shape = st["p"].shape
        P = _capi.params_init(5.0 / 3.0, max(shape) - 2 * NB, 0.0)
        vs = np.sqrt(P.gamma * st["p"] / st["rho"])
        vmax = max((np.abs(st[k]) + vs).max() for k in ("vx", "vy", "vz"))
        hp.timedet(P, vmax, 0.8)
        rc_b, ref, c_b = box_sweep(st, P, d, precision, lim)
        rc_n, out, c_n = noalloc_sweep(st, P, d, precision, lim)
        assert rc_n == rc_b, (name, rc_n, rc_b)
        if lim:
            assert rc_n == _capi.HP_E_NEWTON and c_n.newton_nonconv > 0
        assert (c_n.newton_iters_max, c_n.newton_nonconv, c_n.neg_state_clamps) == \
            (c_b.newton_iters_max, c_b.newton_nonconv, c_b.neg_state_clamps), name
        for k in FIELDS:
            for a, b in zip(faces_and_interior(out[k]), faces_and_interior(ref[k])):
                assert np.array_equal(a, b), (name, k)
