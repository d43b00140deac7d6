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


def diagonal_shock(N, nb=NB):
    """The reference's diagonal 2-D shock (hydropad/initialize_hydro.F90:119-159,
    chooseinit = 4) on one rank: (rho, p) = (1, 1) where the global indices
    satisfy igridx + igridy < N + 1 (igrid = array index + 1, array indices
    nb+1..nb+N), (0.125, 0.1) elsewhere, v = 0."""
    n = N + 2 * nb
    st = empty_state((n, n, n))
    i = np.arange(nb + 1, nb + N + 1)                    # 1-based array indices of the interior
    left = ((i[:, None] + 1) + (i[None, :] + 1)) < N + 1
    sl = tuple(slice(nb, nb + N) for _ in range(3))
    st["rho"][sl] = np.where(left, 1.0, 0.125)[:, :, None]
    st["p"][sl] = np.where(left, 1.0, 0.1)[:, :, None]
    return periodic_fill(st)


def sedov(N, p_blast=100000.0):
    """The reference's Sedov point blast (hydropad/initialize_hydro.F90:161-170,
    chooseinit = 5): rho = p = 1, v = 0, p = 1e5 in the first interior cell."""
    st = uniform(N, rho=1.0, p=1.0)
    st["p"][NB, NB, NB] = p_blast
    return periodic_fill(st)


def isentropic(N, gamma=5.0 / 3.0):
    """Uniform isentropic rest state p = rho^gamma (initialize_hydro.F90:172-179)."""
    return uniform(N, rho=1.0, p=1.0 ** gamma)


def uniform(N, rho=1.0, p=1.0, v=(0.0, 0.0, 0.0)):
    n = N + 2 * NB
    st = empty_state((n, n, n))
    st["rho"][...] = rho
    st["p"][...] = p
    st["vx"][...], st["vy"][...], st["vz"][...] = v
    return st


def smooth_wave(N, ny=3, nz=3, u0=0.5, amp=0.1):
    """One period of a smooth subsonic wave along x in an N x ny x nz box:
    rho = 1 + amp sin(kx), vx = u0 + amp sin(kx + 0.7), p = 1 + amp cos(kx),
    k = 2 pi.  The mean flow keeps |v| > 0, so the kinetic share of the total
    energy is never 0 and an eta1 of 1 sends every cell to the
    internal-energy equation (hydro/integrate.F90:92-98)."""
    st = empty_state((N + 2 * NB, ny + 2 * NB, nz + 2 * NB))
    k = 2.0 * np.pi * cell_centres(N)
    prof = {"rho": 1.0 + amp * np.sin(k), "vx": u0 + amp * np.sin(k + 0.7), "p": 1.0 + amp * np.cos(k)}
    for name, v in prof.items():
        st[name][NB:-NB, NB:-NB, NB:-NB] = v[:, None, None]
    return periodic_fill(st)


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


# ---- comoving boxes ---------------------------------------------------------
# FP32 exponents of the Einstein-de Sitter schedule (hydropad/functions.F90:67-68)
_K23 = float(np.float32(2.0) / np.float32(3.0))
_KM13 = float(np.float32(-1.0) / np.float32(3.0))


def eds_adot(a_start):
    """da/dt at the start of the first step of a box started at a_start, as the
    schedule computes it (initialize_parameters.F90:138-151: t = a^1.5,
    dat = (2/3) t^(-1/3) with the FP32 exponents)."""
    return _K23 * (a_start ** 1.5) ** _KM13


def eds_schedule(a_start, a_end, nsteps):
    """(ca, tstop) such that the Einstein-de Sitter schedule started at a_start
    ends at a = a_end after nsteps steps when the da/a <= ca limit sets dt
    (hydropad/functions.F90:69,87-88): the first step is dt = 0.010 tin
    (initialize_parameters.F90:141), every later one grows a by (1+ca), and the
    last step is clipped at tstop = a_end^(1/K), K the FP32 2/3."""
    tin = a_start ** 1.5
    a1 = (tin * (1.0 + float(np.float32(0.010)))) ** _K23
    ca = (a_end / a1) ** (1.0 / (nsteps - 1)) - 1.0
    return ca * (1.0 + 1e-4), a_end ** (1.0 / _K23)


def grf_box(N, seed, slope=-2.0, rms=0.05, mach=50.0, vfac=None, a_start=1.0 / 51.0, gamma=5.0 / 3.0):
    """Seeded Gaussian random density field with Zel'dovich velocities at
    uniform temperature, on the (N+6)^3 periodic layout.

      delta    P(k) ~ k^slope for 2 pi <= |k| <= pi N (box length 1; the
               single-axis Nyquist planes left empty), zero mean, scaled to
               the given rms; rho = 1 + delta
      v        Zel'dovich growing mode: displacement psi_k = i k delta_k / k^2
               (div psi = -delta) times vfac, by default da/dt at a_start
               (peculiar velocity of linear theory, f = 1 in EdS)
      p        uniform temperature: p = rho cs^2 / gamma with
               cs = v_rms / mach
    The generator is numpy's PCG64 seeded with `seed`; the reference has no
    generator of its own (its initial conditions are read from files,
    hydropad/initialdm.F90:37-70)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((N, N, N))
    wk = np.fft.rfftn(white)
    kx = 2.0 * np.pi * np.fft.fftfreq(N, d=1.0 / N)
    kz = 2.0 * np.pi * np.fft.rfftfreq(N, d=1.0 / N)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k = np.sqrt(k2)
    # the single-axis Nyquist planes carry no derivative (i k X is not
    # Hermitian there), so they are left empty along with |k| beyond pi N
    kmax = np.pi * N * (1.0 - 1e-12)
    band = ((k >= 2.0 * np.pi * (1.0 - 1e-12)) & (k <= np.pi * N * (1.0 + 1e-12))
            & (np.abs(KX) < kmax) & (np.abs(KY) < kmax) & (np.abs(KZ) < kmax))
    amp = np.zeros_like(k)
    amp[band] = k[band] ** (0.5 * slope)
    dk = wk * amp
    delta = np.fft.irfftn(dk, s=(N, N, N), axes=(0, 1, 2))
    f = rms / delta.std()
    dk *= f
    delta *= f
    k2s = np.where(k2 > 0.0, k2, 1.0)
    if vfac is None:
        vfac = eds_adot(a_start)
    vel = [vfac * np.fft.irfftn(1j * K * dk / k2s, s=(N, N, N), axes=(0, 1, 2)) for K in (KX, KY, KZ)]
    rho = 1.0 + delta
    if rho.min() <= 0.0:
        raise ValueError("rms too large: negative density")
    vrms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    cs = vrms / mach
    n = N + 2 * NB
    st = empty_state((n, n, n))
    sl = tuple(slice(NB, NB + N) for _ in range(3))
    st["rho"][sl] = rho
    st["p"][sl] = rho * cs * cs / gamma
    st["vx"][sl], st["vy"][sl], st["vz"][sl] = vel
    return periodic_fill(st)


def boxsize(ngrid):
    """hydropad/init_sys.F90:93-95: each extent over the largest, in FP32."""
    nmax = max(ngrid)
    return tuple(float(np.float32(n) / np.float32(nmax)) for n in ngrid)


def particles(ngrid, npart, seed, vamp=0.3):
    """Seeded particles for the N-body step: positions uniform in the box
    [0, boxsize) (the reference's ppos, hydropad/init_sys.F90:93-95 units),
    velocities normal with deviation vamp.  Returns (ppos, pvel), each
    (npart, 3) C-ordered float64 -- the memory layout of the reference's
    (3, np) arrays (nbody/kick.F90:16-17)."""
    r = np.random.default_rng(seed)
    bs = np.array(boxsize(ngrid))
    ppos = np.ascontiguousarray(r.random((npart, 3)) * bs)
    ppos = np.minimum(ppos, np.nextafter(bs, 0.0))
    pvel = np.ascontiguousarray(vamp * r.standard_normal((npart, 3)))
    return ppos, pvel


def gaussian_gforce(shape, nb=NB):
    """The reference's static gravity set-up, hydropad/initialize_gravity.F90:25-58:
    phi = exp(-|x - c|^2 / (2 sigma^2)) with c = (0.5, 0.5, 0.5), sigma = 0.3
    (a default-REAL literal, so 0.3 rounded to FP32), cell centres
    (real(i) - 0.5) dx taken in FP32 as there; gforce(d) = phi(i+1) - phi(i-1)
    along d (no 1/(2 dx)).  Ghosts of phi and of gforce are periodic images
    (the reference fills them with its halo exchange).  Returns (3, nx, ny, nz)
    Fortran-ordered."""
    nx, ny, nz = shape
    ng = (nx - 2 * nb, ny - 2 * nb, nz - 2 * nb)
    dx = float(np.float32(1.0) / np.float32(max(ng)))
    sigma = float(np.float32(0.3))
    cen = [(np.arange(1, n + 1, dtype=np.float32) - np.float32(0.5)).astype(np.float64) * dx - 0.5 for n in ng]
    r2 = cen[0][:, None, None] ** 2 + cen[1][None, :, None] ** 2 + cen[2][None, None, :] ** 2
    phi = np.zeros(shape, order="F")
    phi[nb:-nb, nb:-nb, nb:-nb] = 1.0 * np.exp(-r2 / (2.0 * sigma ** 2))
    periodic_fill({"phi": phi}, nb)
    g = np.zeros((3,) + tuple(shape), order="F")
    I = slice(nb, -nb)
    g[0][I, I, I] = phi[nb + 1:nx - nb + 1, I, I] - phi[nb - 1:nx - nb - 1, I, I]
    g[1][I, I, I] = phi[I, nb + 1:ny - nb + 1, I] - phi[I, nb - 1:ny - nb - 1, I]
    g[2][I, I, I] = phi[I, I, nb + 1:nz - nb + 1] - phi[I, I, nb - 1:nz - nb - 1]
    for c in range(3):
        a = np.asfortranarray(g[c])
        periodic_fill({"g": a}, nb)
        g[c] = a
    return g
