"""ctypes binding of libhydropad_amd.so (the C-ABI of include/hydropad_gpu.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly -- there is no CPU fallback anywhere in the package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libhydropad_amd.so")

HP_OK, HP_E_ARG, HP_E_HIP, HP_E_NOMEM, HP_E_NEWTON, HP_E_STATE = 0, -1, -2, -3, -4, -5
HP_FP64, HP_MIXED, HP_FP64_EXACT = 0, 1, 2
_ERRNAMES = {HP_E_ARG: "HP_E_ARG", HP_E_HIP: "HP_E_HIP", HP_E_NOMEM: "HP_E_NOMEM",
             HP_E_NEWTON: "HP_E_NEWTON", HP_E_STATE: "HP_E_STATE"}

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_PD = ctypes.POINTER(ctypes.c_double)

PARAM_NAMES = ("gamma", "gamma1", "rgamma", "rgamma1", "m", "rm", "gf",
               "dt", "dx", "rdx", "rdtath", "at", "dat", "rat",
               "flatvalue", "mingradflat", "eta1", "eta2", "dmax")


class HpParams(ctypes.Structure):
    """Mirror of hp_params: the 19 module scalars of hydropad/scalar.F90:98-99."""
    _fields_ = [(n, _D) for n in PARAM_NAMES]

    def copy(self):
        c = HpParams()
        ctypes.pointer(c)[0] = self
        return c

    def as_array(self):
        return np.array([getattr(self, n) for n in PARAM_NAMES], dtype=np.float64)


class HpCounters(ctypes.Structure):
    _fields_ = [("newton_iters_max", ctypes.c_int64), ("newton_nonconv", ctypes.c_int64),
                ("neg_state_clamps", ctypes.c_int64), ("particles_outside", ctypes.c_int64),
                ("reserved", ctypes.c_int64 * 4)]


COSMO_NAMES = ("cour", "ca", "tstop", "t", "dt", "at", "dat", "ath", "dath", "atnew", "datnew", "rat", "rdtath")


class HpCosmo(ctypes.Structure):
    """Mirror of hp_cosmo: the Einstein-de Sitter schedule (hydropad/scalar.F90:16-19)."""
    _fields_ = [(n, _D) for n in COSMO_NAMES] + [("nstep", ctypes.c_int64)]

    def copy(self):
        c = HpCosmo()
        ctypes.pointer(c)[0] = self
        return c


class HydropadError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} failed: {_ERRNAMES.get(code, code)}")
        self.code = code


# every symbol include/hydropad_gpu.h declares (checked by tests/test_capi_symbols.py)
EXPORTS = ("hp_params_init", "hp_create", "hp_destroy", "hp_set_precision", "hp_upload",
           "hp_download", "hp_upload_device", "hp_upload_gforce", "hp_download_device", "hp_set_params",
           "hp_ppmsolver", "hp_exchange_mesh", "hp_ppm_lxlylz", "hp_speed", "hp_evolve",
           "hp_status", "hp_reset_counters", "hp_set_newton_limit", "hp_sync", "hp_integrate_cell", "hp_dualenergy_cell",
           "hp_sweep_timing_enable", "hp_sweep_timing", "hp_set_tiling", "hp_cosmo_init", "hp_timedet_cosmo",
           "hp_cosmo_params", "hp_savetn", "hp_expand", "hp_evolve_cosmo",
           "hp_particles_upload", "hp_particles_download", "hp_kick", "hp_drift",
           "hp_batch_box_bytes", "hp_batch_create", "hp_batch_destroy", "hp_batch_add", "hp_batch_add_sod",
           "hp_sod_of_seed", "hp_batch_remove", "hp_batch_set_newton_limit", "hp_batch_download", "hp_batch_step", "hp_batch_sync",
           "hp_batch_poll", "hp_batch_use_graph", "hp_batch_set_tuning", "hp_batch_get_tuning",
           "hp_batch_bytes_used", "hp_batch_autotune", "hp_serve_run", "hp_ppmsolver_noalloc")

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} not built; run __graft_entry__.build() or hydropad_amd.build.build()")
    L = ctypes.CDLL(LIBPATH)
    PP = ctypes.POINTER(HpParams)
    L.hp_params_init.argtypes = [_D, _I, _D, PP]
    L.hp_create.argtypes = [_I, _I, _I, _I, _I, ctypes.POINTER(_P)]
    L.hp_destroy.argtypes = [_P]
    L.hp_destroy.restype = None
    L.hp_set_precision.argtypes = [_P, _I]
    L.hp_upload.argtypes = [_P] + [_P] * 6
    L.hp_download.argtypes = [_P] + [_P] * 5
    L.hp_upload_device.argtypes = [_P] + [_P] * 6
    L.hp_upload_gforce.argtypes = [_P, _P]
    L.hp_download_device.argtypes = [_P] + [_P] * 5
    L.hp_set_params.argtypes = [_P, PP]
    L.hp_ppmsolver.argtypes = [_P, _I]
    L.hp_exchange_mesh.argtypes = [_P, _I]
    L.hp_ppm_lxlylz.argtypes = [_P, _I]
    L.hp_speed.argtypes = [_P, _PD]
    L.hp_evolve.argtypes = [_P, _I, _I, _D, _D, _PD, _PD]
    L.hp_status.argtypes = [_P, ctypes.POINTER(HpCounters)]
    L.hp_reset_counters.argtypes = [_P]
    L.hp_set_newton_limit.argtypes = [_P, _I]
    L.hp_sync.argtypes = [_P]
    L.hp_integrate_cell.argtypes = [PP] + [_PD] * 18
    L.hp_dualenergy_cell.argtypes = [PP] + [_D] * 8 + [_PD, _PD]
    L.hp_sweep_timing_enable.argtypes = [_P, _I]
    L.hp_sweep_timing.argtypes = [_P, _PD, ctypes.POINTER(ctypes.c_int64)]
    L.hp_set_tiling.argtypes = [_P, _I, _I, _I]
    PC = ctypes.POINTER(HpCosmo)
    L.hp_cosmo_init.argtypes = [_D, _D, _D, _D, PC]
    L.hp_timedet_cosmo.argtypes = [PC, _D, _D]
    L.hp_cosmo_params.argtypes = [PC, PP]
    L.hp_savetn.argtypes = [_P]
    L.hp_expand.argtypes = [_P, _D]
    L.hp_evolve_cosmo.argtypes = [_P, _I, _I, PC]
    L.hp_particles_upload.argtypes = [_P, ctypes.c_int64, _P, _P, _D]
    L.hp_particles_download.argtypes = [_P, _P, _P]
    L.hp_kick.argtypes = [_P, _D]
    L.hp_drift.argtypes = [_P]
    L.hp_ppmsolver_noalloc.argtypes = [PP] + [_P] * 11 + [_I] * 7 + [ctypes.POINTER(HpCounters), _P]
    _lib = L
    return L


def check(fn, rc):
    if rc != HP_OK:
        raise HydropadError(fn, rc)
    return rc


def cosmo_init(a_start, cour, ca, tstop):
    """EdS schedule at a_start (hydropad/initialize_parameters.F90:136-178)."""
    s = HpCosmo()
    check("hp_cosmo_init", lib().hp_cosmo_init(a_start, cour, ca, tstop, ctypes.byref(s)))
    return s


def timedet_cosmo(s, velmax, dx):
    """timedet_cosmo (hydropad/functions.F90:38-151) on the host; advances s in place."""
    check("hp_timedet_cosmo", lib().hp_timedet_cosmo(ctypes.byref(s), velmax, dx))
    return s


def params_init(gamma, ngrid_max, dt=0.0):
    p = HpParams()
    check("hp_params_init", lib().hp_params_init(gamma, int(ngrid_max), dt, ctypes.byref(p)))
    return p
