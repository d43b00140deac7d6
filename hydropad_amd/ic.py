"""Initial conditions for the hydro step, on the host, in the reference layout.

Every generator returns a dict of Fortran-ordered float64 arrays of shape
(N+6)^3 (nbound = 3 ghosts per side) with the ghosts filled periodically, the
state hydropad/init_sys.F90:130-142 hands to the time loop after
initialize_hydro + exchange_mesh.
"""
import numpy as np

NB = 3
FIELDS = ("p", "rho", "vx", "vy", "vz")


def empty_state(shape):
    return {k: np.zeros(shape, dtype=np.float64, order="F") for k in FIELDS}


def periodic_fill(state, nb=NB):
    """Periodic ghost fill, z then x then y (hydropad/exchange_mesh.F90:47-123)."""
    for a in state.values():
        nx, ny, nz = a.shape
        a[:, :, :nb] = a[:, :, nz - 2 * nb:nz - nb]
        a[:, :, nz - nb:] = a[:, :, nb:2 * nb]
        a[:nb, :, :] = a[nx - 2 * nb:nx - nb, :, :]
        a[nx - nb:, :, :] = a[nb:2 * nb, :, :]
        a[:, :nb, :] = a[:, ny - 2 * nb:ny - nb, :]
        a[:, ny - nb:, :] = a[:, nb:2 * nb, :]
    return state


def cell_centres(N):
    """Cell-centre coordinates in [0,1) of the interior cells."""
    return (np.arange(N) + 0.5) / N


def sod(N, axis=0, left=(1.0, 1.0), right=(0.125, 0.1), x0=0.5):
    """Planar Sod tube along `axis`: (rho, p) = left where the cell centre
    x < x0, right otherwise, v = 0 (the 64^3 / 512^3 shock-tube workloads).
    The reference's own split (initialize_hydro.F90:40-47) sits at global cell
    < N/2-nbound-1 instead; the workload definition here is x < 0.5."""
    n = N + 2 * NB
    st = empty_state((n, n, n))
    x = cell_centres(N)
    is_left = x < x0
    shape = [1, 1, 1]
    shape[axis] = N
    m = is_left.reshape(shape)
    sl = tuple(slice(NB, NB + N) for _ in range(3))
    st["rho"][sl] = np.where(m, left[0], right[0])
    st["p"][sl] = np.where(m, left[1], right[1])
    return periodic_fill(st)


def uniform(N, rho=1.0, p=1.0, v=(0.0, 0.0, 0.0)):
    n = N + 2 * NB
    st = empty_state((n, n, n))
    st["rho"][...] = rho
    st["p"][...] = p
    st["vx"][...], st["vy"][...], st["vz"][...] = v
    return st


def random_smooth(N, seed, amp=0.3, vamp=0.3):
    """Seeded random state (rho, p in [1, 1+amp), v in [-vamp/2, vamp/2)):
    cell-to-cell noise that exercises every limiter branch."""
    r = np.random.default_rng(seed)
    n = N + 2 * NB
    st = empty_state((n, n, n))
    sl = tuple(slice(NB, NB + N) for _ in range(3))
    st["rho"][sl] = 1.0 + amp * r.random((N, N, N))
    st["p"][sl] = 1.0 + amp * r.random((N, N, N))
    for k in ("vx", "vy", "vz"):
        st[k][sl] = vamp * (r.random((N, N, N)) - 0.5)
    return periodic_fill(st)


def sod_star(gamma=5.0 / 3.0, left=(1.0, 1.0), right=(0.125, 0.1)):
    """Star state of a Riemann problem with both sides at rest:
    (p*, u*, rho*_L, rho*_R, shock speed), for a left rarefaction and a right
    shock (Toro, Riemann Solvers, ch. 4), solved by bisection to round-off."""
    g = gamma
    rl, pl = left
    rr, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2.0 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-14, 10.0 * max(pl, pr)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid, rl, pl, cl) + f(mid, rr, pr, cr) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    rsl = rl * (ps / pl) ** (1 / g)
    k = (g - 1) / (g + 1)
    rsr = rr * (ps / pr + k) / (k * ps / pr + 1)
    s_shock = cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
    return ps, us, rsl, rsr, s_shock


def sod_exact(x, t, gamma=5.0 / 3.0, left=(1.0, 1.0), right=(0.125, 0.1), x0=0.5):
    """Exact solution (rho, u, p) of the Sod problem at positions x, time t > 0."""
    g = gamma
    rl, pl = left
    rr, pr = right
    cl = np.sqrt(g * pl / rl)
    ps, us, rsl, rsr, s_shock = sod_star(gamma, left, right)
    head, tail = -cl, us - np.sqrt(g * ps / rsl)
    xi = (np.asarray(x, dtype=float) - x0) / t
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    fan_c = 2 / (g + 1) * cl - (g - 1) / (g + 1) * xi
    regions = [xi < head, (xi >= head) & (xi < tail), (xi >= tail) & (xi < us),
               (xi >= us) & (xi < s_shock), xi >= s_shock]
    rho[regions[0]], u[regions[0]], p[regions[0]] = rl, 0.0, pl
    m = regions[1]
    rho[m] = rl * (fan_c[m] / cl) ** (2 / (g - 1))
    u[m] = 2 / (g + 1) * (cl + xi[m])
    p[m] = pl * (fan_c[m] / cl) ** (2 * g / (g - 1))
    rho[regions[2]], u[regions[2]], p[regions[2]] = rsl, us, ps
    rho[regions[3]], u[regions[3]], p[regions[3]] = rsr, us, ps
    rho[regions[4]], u[regions[4]], p[regions[4]] = rr, 0.0, pr
    return rho, u, p
