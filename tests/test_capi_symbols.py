"""The C-ABI library loads without a GPU and exports every entry point that
include/hydropad_gpu.h declares; host-only helpers behave as the reference's
initialisation (no kernel is launched here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hydropad_gpu.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t)\s+(hp_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("hp_ppmsolver", "hp_ppm_lxlylz", "hp_exchange_mesh", "hp_speed", "hp_evolve",
              "hp_integrate_cell", "hp_dualenergy_cell", "hp_create", "hp_upload", "hp_download"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from hydropad_amd import _capi
    if not os.path.exists(_capi.LIBPATH):
        pytest.skip("libhydropad_amd.so not built")
    L = ctypes.CDLL(_capi.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(_capi.EXPORTS)


def test_params_init_matches_reference_initialisation(oracle):
    """hp_params_init is host-only: derived constants of
    initialize_parameters.F90:23-24,64-69 bit for bit against the oracle."""
    from hydropad_amd import _capi
    if not os.path.exists(_capi.LIBPATH):
        pytest.skip("libhydropad_amd.so not built")
    for gamma, n in ((5.0 / 3.0, 64), (1.66666667, 48), (1.4, 192), (5.0 / 3.0, 96)):
        a = _capi.params_init(gamma, n, 0.01).as_array()
        b = oracle.params(gamma=gamma, ngrid=n, dt=0.01).as_array()
        assert np.array_equal(a, b), (gamma, n)
    p = _capi.params_init(5.0 / 3.0, 48, 0.0)
    assert p.dx == 0.020833333954215050


def test_bad_arguments_rejected_without_gpu():
    from hydropad_amd import _capi
    if not os.path.exists(_capi.LIBPATH):
        pytest.skip("libhydropad_amd.so not built")
    L = _capi.lib()
    p = _capi.HpParams()
    assert L.hp_params_init(5.0 / 3.0, 0, 0.0, ctypes.byref(p)) == _capi.HP_E_ARG
    assert L.hp_params_init(1.0, 64, 0.0, ctypes.byref(p)) == _capi.HP_E_ARG
    h = ctypes.c_void_p()
    assert L.hp_create(70, 70, 70, 2, 0, ctypes.byref(h)) == _capi.HP_E_ARG     # nbound must be 3
    assert L.hp_create(70, 70, 70, 3, 7, ctypes.byref(h)) == _capi.HP_E_ARG     # unknown precision
    assert L.hp_create(1300, 1300, 1300, 3, 0, ctypes.byref(h)) == _capi.HP_E_ARG   # past 32-bit cell indexing
    assert L.hp_ppmsolver(None, 1) == _capi.HP_E_ARG


def test_batch_host_helpers_without_gpu():
    """Host-only pieces of the batch API: memory accounting, the seed -> Sod
    mapping, and argument checks that return before any HIP call."""
    from hydropad_amd import _capi, batch
    if not os.path.exists(_capi.LIBPATH):
        pytest.skip("libhydropad_amd.so not built")
    fb = -(-70 ** 3 * 8 // 256) * 256
    assert batch.box_bytes((70, 70, 70)) == 10 * fb
    assert batch.box_bytes((70, 70, 70), comoving=True) == 15 * fb
    seen_axis, seen_flip, xs = set(), set(), []
    for seed in range(300):
        a, x0, f = batch.sod_of_seed(seed)
        assert a in (0, 1, 2) and f in (0, 1) and 0.25 <= x0 < 0.75
        seen_axis.add(a)
        seen_flip.add(f)
        xs.append(x0)
    assert seen_axis == {0, 1, 2} and seen_flip == {0, 1}
    assert np.std(xs) > 0.1
    assert batch.sod_of_seed(12345) == batch.sod_of_seed(12345)
    L = batch._lib()
    h = ctypes.c_void_p()
    assert L.hp_batch_create(0, 1 << 30, ctypes.byref(h)) == _capi.HP_E_ARG
    assert L.hp_batch_create(100000, 1 << 30, ctypes.byref(h)) == _capi.HP_E_ARG
    assert L.hp_batch_step(None, 1) == _capi.HP_E_ARG
    assert L.hp_serve_run(None, 0, 0, 1, 0, None, None, None) == _capi.HP_E_ARG
