"""Pin the oracle's internal-energy branch of integrate (e/E < eta1), where the
new pressure comes from the internal-energy equation with
pmean = (pm1+pm2)/2 (hydro/ppm.F90:583; undefined in the reference's active
hydro/integrate.F90:19,60-75), by what the mathematics fixes:

* Convergence.  The same smooth subsonic wave run with the dual-energy switch
  forced each way -- eta1 = 0 (p from the total energy everywhere) and
  eta1 = 1, eta2 = 2 (p from the internal-energy equation everywhere) -- are
  two consistent second-order discretisations of one smooth flow, so their
  difference in p must fall about 4x per doubling of N.  A p dv/dx term
  taken at one face (pmean = pm1) is first order (2x), and a missing 1/2 or
  a flipped sign does not converge at all (1x): each is planted in the
  oracle below and must fail the same check.
* Closed forms of the cell update: with equal fluxes on both faces the
  internal-energy equation leaves the caller's eint unchanged, so
  pnew = (gamma-1) eint exactly in that regime, and the e <= 0 fallback
  keeps the old pressure (integrate.F90:75-81,92-105).  The reference's own
  build cannot be compared on the pressure there: its pmean is an
  uninitialised local, and the flang build of oracle/_ref returns a NaN
  pressure in this regime even though pmean only multiplies a zero velocity
  jump.  Its density and velocities must still agree bit for bit.
"""
import numpy as np
import pytest

from helpers import NB
from hydropad_amd import ic
from oracle import Reference

T_END = 0.1
COUR = 0.4


def run_wave(oracle, N, eta1, eta2):
    """The wave to t = 0.1 with a fixed dt (CFL 0.4 on max |u|+c < 2)."""
    st = ic.smooth_wave(N)
    P = oracle.params(ngrid=N)
    P.eta1, P.eta2 = eta1, eta2
    nsteps = int(np.ceil(T_END / (COUR * P.dx / 2.0) - 1e-9))
    P.dt = P.rdtath = T_END / nsteps
    for n in range(1, nsteps + 1):
        assert oracle.ppm_lxlylz(P, st, 1 if n == 1 else n % 6 + 1) >= 0
    return st


def form_difference(oracle, N):
    a = run_wave(oracle, N, 0.0, 0.1)     # pressure from the total energy
    b = run_wave(oracle, N, 1.0, 2.0)     # pressure from the internal-energy equation
    return float(np.abs(a["p"][NB:-NB, NB, NB] - b["p"][NB:-NB, NB, NB]).mean())


def ratios(oracle):
    d = [form_difference(oracle, N) for N in (16, 32, 64)]
    return d, [d[0] / d[1], d[1] / d[2]]


def test_internal_energy_branch_converges_second_order(oracle):
    d, r = ratios(oracle)
    assert all(x >= 3.0 for x in r), (d, r)
    assert d[2] < 2e-6, d


@pytest.mark.parametrize("what", [1, 2, 3], ids=["pmean=pm1", "pmean=pm1+pm2", "pmean=-mean"])
def test_planted_pmean_mistakes_fail_the_convergence_check(oracle, what):
    oracle.plant(what)
    try:
        d, r = ratios(oracle)
    finally:
        oracle.plant(0)
    assert min(r) < 2.5, (what, d, r)


def test_wave_regimes(oracle):
    """eta1 = 1 really puts every cell on the internal-energy equation: the
    kinetic share of the total energy is positive everywhere."""
    st = ic.smooth_wave(32)
    sl = (slice(NB, -NB),) * 3
    e = st["p"][sl] * 1.5
    E = e + 0.5 * st["rho"][sl] * st["vx"][sl] ** 2
    ee = e / E
    assert ee.max() < 1.0 and ee.min() > 0.8


def cold_cell(eint):
    """Equal fluxes on both faces, a cold fast cell (e/E ~ 1e-4 < eta1) and a
    caller eint unrelated to p/(gamma-1)."""
    fl = [0.3, 0.3, 1.1, 1.1, 2.0, 2.0, 0.1, 0.1, -0.2, -0.2]
    pres, rho, vx = 1e-4, 1.0, 2.0
    st = [pres, rho, vx, 0.0, 0.0]
    etot = pres * 1.5 + 0.5 * rho * vx * vx
    return fl, st, etot, eint


@pytest.mark.parametrize("eint", [3e-4, 7.5e-5])
def test_integrate_uses_callers_eint(oracle, eint):
    P = oracle.params(ngrid=64, dt=0.01)
    fl, st, etot, e = cold_cell(eint)
    out = oracle.integrate(P, fl, st, etot, e, 0.0)
    assert out[0] == (P.gamma - 1) * eint
    assert out[1] == st[1]
    if Reference.available():
        R = Reference()
        R.set_params(P)
        assert np.array_equal(R.integrate(fl, st, etot, e, 0.0)[1:], out[1:])


@pytest.mark.parametrize("eint", [0.0, -1.0])
def test_integrate_nonpositive_eint_keeps_pressure(oracle, eint):
    P = oracle.params(ngrid=64, dt=0.01)
    fl, st, etot, e = cold_cell(eint)
    out = oracle.integrate(P, fl, st, etot, e, 0.0)
    assert out[0] == st[0]
    if Reference.available():
        R = Reference()
        R.set_params(P)
        assert np.array_equal(R.integrate(fl, st, etot, e, 0.0)[1:], out[1:])
