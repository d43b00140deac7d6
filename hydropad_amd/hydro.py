"""Host-side mirror of Hydropad's hydro-step interface over the C-ABI.

`Box` owns one device-resident periodic box (an hp_ctx).  Its methods carry
the reference's subroutine names and argument meaning:

    ppmsolver(ndirection)   hydro/ppmsolver.F90:1-5      one sweep, in place
    ppm_lxlylz(nrot)        hydro/ppm_lxlylz.F90:1-223   three sweeps in nrot order
    exchange_mesh(dir=0)    hydropad/exchange_mesh.F90   periodic ghost fill
    speed()                 hydropad/speed.F90:2-59      CFL maximum velocity
    timedet(velmax, cour)   hydropad/functions.F90:2-35  next dt
    evolve(first, n, cour)  hydropad/main.F90:100-121 + evolve_sys.F90

Fields are Fortran-ordered float64 arrays of shape (nx, ny, nz) including
nbound=3 ghost cells per side, as the reference stores p3d/rho3d/vx3d/vy3d/vz3d.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import HP_FP64, HP_FP64_EXACT, HP_MIXED, HpCounters, HpParams, check

FIELDS = ("p", "rho", "vx", "vy", "vz")
NBOUND = 3
PRECISIONS = {"fp64": HP_FP64, "mixed": HP_MIXED, "fp64_exact": HP_FP64_EXACT}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nrot_of_step(nstep):
    """Sweep-order index of step nstep (hydropad/evolve_sys.F90:30-31)."""
    return 1 if nstep == 1 else nstep % 6 + 1


NROT_ORDER = {1: (1, 2, 3), 2: (3, 2, 1), 3: (2, 3, 1), 4: (1, 3, 2), 5: (3, 1, 2), 6: (2, 1, 3)}


def timedet(params, velmax, cour, tstop=1e30, t=0.0):
    """Non-cosmological timedet (hydropad/functions.F90:2-35) on the host.

    Returns (dt, t_new) and updates params.dt / params.rdtath in place; the
    arithmetic is the reference's, in IEEE double."""
    told = t
    rdtc = velmax / (cour * params.dx)
    dt = 0.1 if velmax == 0.0 else 1.0 / rdtc
    t = told + dt
    if t > tstop:
        t = tstop
        dt = t - told
    params.at, params.dat, params.rat = 1.0, 0.0, 1.0
    params.dt = dt
    params.rdtath = dt / 1.0
    return dt, t


class Box:
    """One periodic box resident on the GPU."""

    def __init__(self, shape, precision="fp64", gamma=5.0 / 3.0, nbound=NBOUND):
        if isinstance(shape, int):
            shape = (shape + 2 * nbound,) * 3
        self.shape = tuple(int(s) for s in shape)
        self.nbound = nbound
        self.interior = tuple(s - 2 * nbound for s in self.shape)
        self.L = _capi.lib()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = ctypes.c_void_p()
        check("hp_create", self.L.hp_create(*self.shape, nbound, prec, ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self.params = _capi.params_init(gamma, max(self.interior), 0.0)
        self.set_params(self.params)

    # ---- lifetime ----
    def close(self):
        if self._h:
            self.L.hp_destroy(self._h)
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

    # ---- state ----
    def upload(self, state, gforce=None):
        arrs = []
        for k in FIELDS:
            a = state[k]
            if a.shape != self.shape or a.dtype != np.float64 or not a.flags.f_contiguous:
                raise ValueError(f"field {k} must be float64 Fortran-ordered {self.shape}")
            arrs.append(a)
        g = None
        if gforce is not None:
            g = np.asfortranarray(gforce, dtype=np.float64)
            if g.shape != (3,) + self.shape:
                raise ValueError("gforce must be (3, nx, ny, nz)")
        check("hp_upload", self.L.hp_upload(self._h, *[_ptr(a) for a in arrs],
                                             _ptr(g) if g is not None else None))

    def upload_gforce(self, gforce):
        """The gravity array alone ((3, nx, ny, nz) or None for none); the
        fields stay as they are."""
        g = None
        if gforce is not None:
            g = np.asfortranarray(gforce, dtype=np.float64)
            if g.shape != (3,) + self.shape:
                raise ValueError("gforce must be (3, nx, ny, nz)")
        check("hp_upload_gforce", self.L.hp_upload_gforce(self._h, _ptr(g) if g is not None else None))

    def download(self):
        out = {k: np.empty(self.shape, dtype=np.float64, order="F") for k in FIELDS}
        check("hp_download", self.L.hp_download(self._h, *[_ptr(out[k]) for k in FIELDS]))
        return out

    def set_params(self, params):
        self.params = params
        check("hp_set_params", self.L.hp_set_params(self._h, ctypes.byref(params)))

    def set_precision(self, precision):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check("hp_set_precision", self.L.hp_set_precision(self._h, prec))
        self.precision = precision

    # ---- the reference's subroutines ----
    def ppmsolver(self, ndirection):
        check("hp_ppmsolver", self.L.hp_ppmsolver(self._h, int(ndirection)))

    def exchange_mesh(self, direction=0):
        check("hp_exchange_mesh", self.L.hp_exchange_mesh(self._h, int(direction)))

    def ppm_lxlylz(self, nrot):
        check("hp_ppm_lxlylz", self.L.hp_ppm_lxlylz(self._h, int(nrot)))

    def speed(self):
        v = ctypes.c_double()
        check("hp_speed", self.L.hp_speed(self._h, ctypes.byref(v)))
        return v.value

    def timedet(self, velmax, cour, tstop=1e30, t=0.0):
        dt, t = timedet(self.params, velmax, cour, tstop, t)
        self.set_params(self.params)
        return dt, t

    def evolve(self, first_step, nsteps, cour, t, dt, tstop=1e30):
        """Device-resident time loop; returns (t, dt) after the last step."""
        tt, dd = ctypes.c_double(t), ctypes.c_double(dt)
        check("hp_evolve", self.L.hp_evolve(self._h, int(first_step), int(nsteps), cour, tstop,
                                            ctypes.byref(tt), ctypes.byref(dd)))
        self.params.dt = dd.value
        self.params.rdtath = dd.value
        return tt.value, dd.value

    # ---- particles (nbody/) ----
    def particles_upload(self, ppos, pvel, dtat):
        """The reference's ppos/pvel as (np, 3) C-ordered float64 arrays (the
        memory of its (3, np) arrays) and the kicks' time step dtat."""
        arrs = []
        for a in (ppos, pvel):
            if a.ndim != 2 or a.shape[1] != 3 or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("particle arrays must be float64 (np, 3) C-contiguous")
            arrs.append(a)
        if ppos.shape != pvel.shape:
            raise ValueError("ppos and pvel differ in shape")
        check("hp_particles_upload", self.L.hp_particles_upload(self._h, ppos.shape[0], _ptr(arrs[0]),
                                                                _ptr(arrs[1]), float(dtat)))
        self.npart = ppos.shape[0]

    def particles_download(self):
        n = getattr(self, "npart", 0)
        x, v = np.empty((n, 3)), np.empty((n, 3))
        check("hp_particles_download", self.L.hp_particles_download(self._h, _ptr(x), _ptr(v)))
        return x, v

    def kick(self, dtat):
        """SUBROUTINE kick (nbody/kick.F90) on the resident particles."""
        check("hp_kick", self.L.hp_kick(self._h, float(dtat)))

    def drift(self):
        """SUBROUTINE drift (nbody/drift.F90) with the current rdtath."""
        check("hp_drift", self.L.hp_drift(self._h))

    # ---- comoving boxes ----
    def set_cosmo(self, sched):
        """Take the schedule's at, dat, rat, dt, rdtath as the sweep scalars."""
        check("hp_cosmo_params", self.L.hp_cosmo_params(ctypes.byref(sched), ctypes.byref(self.params)))
        self.set_params(self.params)

    def savetn(self):
        check("hp_savetn", self.L.hp_savetn(self._h))

    def expand(self, afact):
        check("hp_expand", self.L.hp_expand(self._h, afact))

    def evolve_cosmo(self, first_step, nsteps, sched):
        """Device-resident comoving time loop; advances sched (HpCosmo) in place."""
        check("hp_evolve_cosmo", self.L.hp_evolve_cosmo(self._h, int(first_step), int(nsteps),
                                                        ctypes.byref(sched)))
        check("hp_cosmo_params", self.L.hp_cosmo_params(ctypes.byref(sched), ctypes.byref(self.params)))
        return sched

    def set_tiling(self, x_chunks=0, yz_seglen=0, block=0):
        """Launch shape of the sweeps (never changes results)."""
        check("hp_set_tiling", self.L.hp_set_tiling(self._h, int(x_chunks), int(yz_seglen), int(block)))

    def sweep_timing_enable(self, on=True):
        check("hp_sweep_timing_enable", self.L.hp_sweep_timing_enable(self._h, 1 if on else 0))

    def sweep_timing(self):
        """(ms[3], launches[3]) summed per sweep direction since enabling."""
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_int64 * 3)()
        check("hp_sweep_timing", self.L.hp_sweep_timing(self._h, ms, n))
        return list(ms), list(n)

    def counters(self):
        c = HpCounters()
        rc = self.L.hp_status(self._h, ctypes.byref(c))
        if rc not in (_capi.HP_OK, _capi.HP_E_NEWTON):
            check("hp_status", rc)
        return c

    def reset_counters(self):
        check("hp_reset_counters", self.L.hp_reset_counters(self._h))

    def set_newton_limit(self, nmax=0):
        """Iterations after which a Riemann Newton iteration counts as not
        converged (hydro/riemann.F90:51, nnewtmax = 100; 0 restores 100)."""
        check("hp_set_newton_limit", self.L.hp_set_newton_limit(self._h, int(nmax)))

    def sync(self):
        check("hp_sync", self.L.hp_sync(self._h))


def integrate_cell(params, fl10, st5, etot, eint, ghalf):
    """hydro/integrate.F90 on one cell, run on the device; returns new (p,rho,vx,vy,vz)."""
    L = _capi.lib()
    a = [ctypes.c_double(float(v)) for v in list(fl10) + list(st5) + [etot, eint, ghalf]]
    check("hp_integrate_cell", L.hp_integrate_cell(ctypes.byref(params), *[ctypes.byref(v) for v in a]))
    return np.array([a[10 + i].value for i in range(5)])


def dualenergy_cell(params, etotnew, rhonew, vx, vy, vz, eintnew_in, eint_old, pres_old):
    """Dual-energy selection of hydro/integrate.F90:196-216, run on the device."""
    L = _capi.lib()
    pn, en = ctypes.c_double(), ctypes.c_double()
    check("hp_dualenergy_cell", L.hp_dualenergy_cell(ctypes.byref(params), etotnew, rhonew, vx, vy, vz,
                                                     eintnew_in, eint_old, pres_old, ctypes.byref(pn),
                                                     ctypes.byref(en)))
    return pn.value, en.value
