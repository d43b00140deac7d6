"""CPU oracle for the PPM hydro path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product path (hydropad_amd) never does.

`Oracle` binds liboracle.so (hp_oracle.c, the C restatement of the
reference's STENCIL path).  `Reference` binds _ref/libhydropad_ref.so, the
reference's own Fortran routines compiled from its sources by the Makefile
here, when that build exists.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libhydropad_ref.so")

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p

PARAM_NAMES = ("gamma", "gamma1", "rgamma", "rgamma1", "m", "rm", "gf",
               "dt", "dx", "rdx", "rdtath", "at", "dat", "rat",
               "flatvalue", "mingradflat", "eta1", "eta2", "dmax")


class HpParams(ctypes.Structure):
    _fields_ = [(n, _D) for n in PARAM_NAMES]

    def as_array(self):
        return np.array([getattr(self, n) for n in PARAM_NAMES], dtype=np.float64)

    def copy(self):
        c = HpParams()
        ctypes.pointer(c)[0] = self
        return c


class HpCounters(ctypes.Structure):
    _fields_ = [("newton_iters_max", ctypes.c_int64), ("newton_nonconv", ctypes.c_int64),
                ("neg_state_clamps", ctypes.c_int64), ("particles_outside", ctypes.c_int64),
                ("reserved", ctypes.c_int64 * 4)]


COSMO_NAMES = ("cour", "ca", "tstop", "t", "dt", "at", "dat", "ath", "dath", "atnew", "datnew", "rat", "rdtath")


class HpCosmo(ctypes.Structure):
    """Mirror of hp_cosmo (the Einstein-de Sitter schedule state)."""
    _fields_ = [(n, _D) for n in COSMO_NAMES] + [("nstep", ctypes.c_int64)]

    def copy(self):
        c = HpCosmo()
        ctypes.pointer(c)[0] = self
        return c


def build(ref=True):
    """Compile liboracle.so (and the reference build when its tree exists)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """ctypes binding of liboracle.so."""

    def __init__(self, path=LIB):
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        self.L = L
        PP = ctypes.POINTER(HpParams)
        L.or_params_init.argtypes = [_D, _I, _D, PP]
        L.or_order2interp.argtypes = [PP] + [_P] * 8 + [ctypes.POINTER(ctypes.c_int64)]
        L.or_riemannsolver.argtypes = [PP, _P, _P, _P, _P]
        L.or_riemannsolver.restype = _I
        L.or_riemann.argtypes = [PP, ctypes.POINTER(_D), ctypes.POINTER(_D)] + [_D] * 6
        L.or_riemann.restype = _I
        L.or_integrate.argtypes = [PP] + [_D] * 10 + [ctypes.POINTER(_D)] * 5 + [_D] * 3
        L.or_internal.argtypes = [PP, _D]
        L.or_internal.restype = _D
        L.or_total.argtypes = [PP] + [_D] * 5
        L.or_total.restype = _D
        L.or_ppmsolver.argtypes = [PP] + [_P] * 6 + [_I] * 5 + [ctypes.POINTER(HpCounters)]
        L.or_ppmsolver.restype = _I
        L.or_exchange_mesh.argtypes = [_P] * 5 + [_I] * 5
        L.or_speed.argtypes = [PP] + [_P] * 5 + [_I] * 4
        L.or_speed.restype = _D
        L.or_timedet.argtypes = [PP, _D, _D, _D, ctypes.POINTER(_D)]
        L.or_timedet.restype = _D
        L.or_ppm_lxlylz.argtypes = [PP] + [_P] * 6 + [_I] * 5 + [ctypes.POINTER(HpCounters)]
        L.or_ppm_lxlylz.restype = _I
        L.or_evolve.argtypes = [PP] + [_P] * 6 + [_I] * 6 + [_D, _D, ctypes.POINTER(_D),
                                                              ctypes.POINTER(HpCounters)]
        L.or_evolve.restype = _I
        PC = ctypes.POINTER(HpCosmo)
        L.or_cosmo_init.argtypes = [_D, _D, _D, _D, PC]
        L.or_cosmo_init.restype = None
        L.or_timedet_cosmo.argtypes = [PC, _D, _D]
        L.or_timedet_cosmo.restype = None
        L.or_expand.argtypes = [_D] + [_P] * 8 + [ctypes.c_long]
        L.or_expand.restype = None
        L.or_evolve_cosmo.argtypes = [PP] + [_P] * 6 + [_I] * 6 + [PC, ctypes.POINTER(HpCounters)]
        L.or_evolve_cosmo.restype = _I
        L.or_kick.argtypes = [PP, ctypes.c_int64, _P, _P, _P, _I, _I, _I, _I, _D]
        L.or_kick.restype = ctypes.c_int64
        L.or_drift.argtypes = [PP, ctypes.c_int64, _P, _P, _I, _I, _I]
        L.or_drift.restype = None
        L.or_particle_vmax.argtypes = [ctypes.c_int64, _P]
        L.or_particle_vmax.restype = _D
        L.or_set_threads.argtypes = [_I]
        L.or_set_newton_max.argtypes = [_I]
        L.or_plant.argtypes = [_I]
        L.or_nrot_dir.argtypes = [_I, _I]
        L.or_nrot_dir.restype = _I

    # ---- N-body: ppos, pvel are (3, np) Fortran-ordered (np, 3 C-contiguous) ----
    def kick(self, P, ppos, pvel, gforce, dtat, nbound=3):
        """nbody/kick.F90 in place on pvel; returns the particles skipped
        because their CIC cube leaves the array."""
        g = np.asfortranarray(gforce, dtype=np.float64)
        _, nx, ny, nz = g.shape
        return self.L.or_kick(ctypes.byref(P), ppos.shape[0], _ptr(ppos), _ptr(pvel), _ptr(g), nx, ny, nz,
                              nbound, dtat)

    def drift(self, P, ppos, pvel, ngrid):
        """nbody/drift.F90 in place on ppos; ngrid = (ngridx, ngridy, ngridz)."""
        self.L.or_drift(ctypes.byref(P), ppos.shape[0], _ptr(ppos), _ptr(pvel), *ngrid)

    def particle_vmax(self, pvel):
        return self.L.or_particle_vmax(pvel.shape[0], _ptr(pvel))

    def set_threads(self, n):
        self.L.or_set_threads(int(n))

    def plant(self, what=0):
        """Plant a known mistake (tests that show a check can fail): 1 pmean =
        pm1, 2 pmean = pm1 + pm2, 3 pmean = -(pm1 + pm2)/2; 0 removes it."""
        self.L.or_plant(int(what))

    def set_newton_max(self, n=0):
        """Newton iteration limit (hydro/riemann.F90:51 has 100; 0 restores it)."""
        self.L.or_set_newton_max(int(n))

    def params(self, gamma=5.0 / 3.0, ngrid=64, dt=0.0):
        p = HpParams()
        self.L.or_params_init(gamma, ngrid, dt, ctypes.byref(p))
        return p

    def order2interp(self, P, pres, rho, vx, vy, vz, g, c):
        out = np.zeros(10)
        arrs = [_f64(a) for a in (pres, rho, vx, vy, vz, g, c)]
        n = ctypes.c_int64(0)
        self.L.or_order2interp(ctypes.byref(P), *[_ptr(a) for a in arrs], _ptr(out), ctypes.byref(n))
        return out, n.value

    def riemannsolver(self, P, rhobar, vxbar, state10):
        out = np.zeros(5)
        rb, vb, s = _f64(rhobar), _f64(vxbar), _f64(state10)
        nit = self.L.or_riemannsolver(ctypes.byref(P), _ptr(rb), _ptr(vb), _ptr(s), _ptr(out))
        return out, nit

    def riemann(self, P, pguess, pl, pr, rl, rr, vl, vr):
        ps, vs = _D(pguess), _D(0.0)
        nit = self.L.or_riemann(ctypes.byref(P), ctypes.byref(ps), ctypes.byref(vs), pl, pr, rl, rr, vl, vr)
        return ps.value, vs.value, nit

    def integrate(self, P, fl10, st5, etot, eint, ghalf):
        st = [_D(v) for v in st5]
        self.L.or_integrate(ctypes.byref(P), *[float(v) for v in fl10],
                            *[ctypes.byref(v) for v in st], etot, eint, ghalf)
        return np.array([v.value for v in st])

    def internal(self, P, p):
        return self.L.or_internal(ctypes.byref(P), p)

    def total(self, P, p, rho, vx, vy, vz):
        return self.L.or_total(ctypes.byref(P), p, rho, vx, vy, vz)

    # ---- whole-box operations on Fortran-ordered (nx,ny,nz) arrays ----
    @staticmethod
    def _fields(state):
        for k in ("p", "rho", "vx", "vy", "vz"):
            a = state[k]
            assert a.dtype == np.float64 and a.flags.f_contiguous, k
        return [state[k] for k in ("p", "rho", "vx", "vy", "vz")]

    def ppmsolver(self, P, state, ndir, nbound=3, gforce=None, counters=None):
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        cnt = counters if counters is not None else HpCounters()
        g = None if gforce is None else np.asfortranarray(gforce, dtype=np.float64)
        return self.L.or_ppmsolver(ctypes.byref(P), *[_ptr(a) for a in f], _ptr(g), nx, ny, nz, nbound,
                                   ndir, ctypes.byref(cnt))

    def exchange_mesh(self, state, nbound=3, which=0):
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        self.L.or_exchange_mesh(*[_ptr(a) for a in f], nx, ny, nz, nbound, which)

    def speed(self, P, state, nbound=3):
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        return self.L.or_speed(ctypes.byref(P), *[_ptr(a) for a in f], nx, ny, nz, nbound)

    def timedet(self, P, velmax, cour, tstop, t):
        tt = _D(t)
        dt = self.L.or_timedet(ctypes.byref(P), velmax, cour, tstop, ctypes.byref(tt))
        return dt, tt.value

    def ppm_lxlylz(self, P, state, nrot, nbound=3, gforce=None, counters=None):
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        cnt = counters if counters is not None else HpCounters()
        g = None if gforce is None else np.asfortranarray(gforce, dtype=np.float64)
        return self.L.or_ppm_lxlylz(ctypes.byref(P), *[_ptr(a) for a in f], _ptr(g), nx, ny, nz, nbound,
                                    nrot, ctypes.byref(cnt))

    def evolve(self, P, state, first_step, nsteps, cour, tstop=1e30, t=0.0, nbound=3, gforce=None,
               counters=None):
        """Time loop; P.dt/P.rdtath must already hold the first step's dt."""
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        cnt = counters if counters is not None else HpCounters()
        tt = _D(t)
        g = None if gforce is None else np.asfortranarray(gforce, dtype=np.float64)
        r = self.L.or_evolve(ctypes.byref(P), *[_ptr(a) for a in f], _ptr(g), nx, ny, nz, nbound,
                             first_step, nsteps, cour, tstop, ctypes.byref(tt), ctypes.byref(cnt))
        return r, tt.value

    def nrot_dir(self, nrot, s):
        return self.L.or_nrot_dir(nrot, s)

    # ---- comoving boxes ----
    def cosmo_init(self, a_start, cour, ca, tstop):
        s = HpCosmo()
        self.L.or_cosmo_init(a_start, cour, ca, tstop, ctypes.byref(s))
        return s

    def timedet_cosmo(self, s, velmax, dx):
        self.L.or_timedet_cosmo(ctypes.byref(s), velmax, dx)
        return s

    def expand(self, afact, state, old):
        f = self._fields(state)
        o = [np.asfortranarray(old[k], dtype=np.float64) for k in ("p", "vx", "vy", "vz")]
        self.L.or_expand(afact, _ptr(f[0]), _ptr(f[2]), _ptr(f[3]), _ptr(f[4]), *[_ptr(a) for a in o], f[0].size)

    def evolve_cosmo(self, P, state, first_step, nsteps, sched, nbound=3, gforce=None, counters=None):
        """Comoving time loop; sched (HpCosmo) is advanced in place, P takes its scalars."""
        f = self._fields(state)
        nx, ny, nz = f[0].shape
        cnt = counters if counters is not None else HpCounters()
        g = None if gforce is None else np.asfortranarray(gforce, dtype=np.float64)
        return self.L.or_evolve_cosmo(ctypes.byref(P), *[_ptr(a) for a in f], _ptr(g), nx, ny, nz, nbound,
                                      first_step, nsteps, ctypes.byref(sched), ctypes.byref(cnt))


class Reference:
    """ctypes binding of the reference's own routines (oracle/_ref)."""

    def __init__(self, path=REFLIB):
        L = ctypes.CDLL(path)
        self.L = L
        L.ref_set_params.argtypes = [_P]
        L.ref_ppmsolver.argtypes = [_P] * 6 + [_I] * 5
        L.ref_order2interp.argtypes = [_P] * 8
        L.ref_riemannsolver.argtypes = [_P] * 4
        L.ref_riemann.argtypes = [_P] + [_D] * 6
        L.ref_integrate.argtypes = [_P] * 3
        L.ref_nbody_setup.argtypes = [_I] * 4
        L.ref_kick.argtypes = [_I, _P, _P, _I, _I, _I, _P, _D, _D]
        L.ref_drift.argtypes = [_I, _P, _P, _D]

    @staticmethod
    def available():
        return os.path.exists(REFLIB)

    def set_params(self, P):
        self._prm = P.as_array()
        self.L.ref_set_params(_ptr(self._prm))

    def order2interp(self, pres, rho, vx, vy, vz, g, c):
        out = np.zeros(10)
        arrs = [_f64(a) for a in (pres, rho, vx, vy, vz, g, c)]
        self.L.ref_order2interp(*[_ptr(a) for a in arrs], _ptr(out))
        return out

    def riemannsolver(self, rhobar, vxbar, state10):
        out = np.zeros(5)
        rb, vb, s = _f64(rhobar), _f64(vxbar), _f64(state10)
        self.L.ref_riemannsolver(_ptr(rb), _ptr(vb), _ptr(s), _ptr(out))
        return out

    def riemann(self, pguess, pl, pr, rl, rr, vl, vr):
        io = np.array([pguess, 0.0])
        self.L.ref_riemann(_ptr(io), pl, pr, rl, rr, vl, vr)
        return io[0], io[1]

    def integrate(self, fl10, st5, etot, eint, ghalf):
        f = _f64(fl10)
        st = np.array(st5, dtype=np.float64)
        e = np.array([etot, eint, ghalf])
        self.L.ref_integrate(_ptr(f), _ptr(st), _ptr(e))
        return st

    def kick(self, ppos, pvel, gforce, dat, dtat, nbound=3):
        g = np.asfortranarray(gforce, dtype=np.float64)
        _, nx, ny, nz = g.shape
        self.L.ref_nbody_setup(nx - 2 * nbound, ny - 2 * nbound, nz - 2 * nbound, nbound)
        self.L.ref_kick(ppos.shape[0], _ptr(ppos), _ptr(pvel), nx, ny, nz, _ptr(g), dat, dtat)

    def drift(self, ppos, pvel, rdtath, ngrid, nbound=3):
        self.L.ref_nbody_setup(*ngrid, nbound)
        self.L.ref_drift(ppos.shape[0], _ptr(ppos), _ptr(pvel), rdtath)

    def ppmsolver(self, state, ndir, nbound=3, gforce=None):
        f = Oracle._fields(state)
        nx, ny, nz = f[0].shape
        if gforce is None:
            gforce = np.zeros((3, nx, ny, nz), order="F")
        g = np.asfortranarray(gforce, dtype=np.float64)
        self.L.ref_ppmsolver(*[_ptr(a) for a in f], _ptr(g), nx, ny, nz, nbound, ndir)
