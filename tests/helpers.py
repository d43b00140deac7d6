"""Shared helpers of the test suite (no GPU, no oracle imports here)."""
import numpy as np

FIELDS = ("p", "rho", "vx", "vy", "vz")
NB = 3


def copy_state(st):
    return {k: v.copy(order="F") for k, v in st.items()}


def interior(a, nb=NB):
    return a[nb:-nb, nb:-nb, nb:-nb]


def rel_l1(a, b):
    """sum|a-b| / sum|b| over the interior."""
    a, b = interior(a), interior(b)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def rel_linf(a, b):
    a, b = interior(a), interior(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def state_from_stack(a):
    return {k: np.asfortranarray(a[i]) for i, k in enumerate(FIELDS)}


def to_hp_params(src, cls):
    p = cls()
    for name, _ in cls._fields_:
        setattr(p, name, float(getattr(src, name)))
    return p


def riemann_case(P, s):
    """Wave pattern riemannsolver.F90:61-64 selects for the interface state
    s = (pl, pr, rhol, rhor, vl, vr, ...): 1 two shocks, 2 shock +
    rarefaction, 3 two rarefactions, 4 vacuum."""
    pl, pr, rl, rr, vl, vr = s[:6]
    cl, cr = np.sqrt(P.gamma * pl / rl), np.sqrt(P.gamma * pr / rr)
    p1, p2, r1, c1, c2 = (pl, pr, rl, cl, cr) if pl <= pr else (pr, pl, rr, cr, cl)
    dvv = vl - vr
    dv2s = (p2 - p1) / np.sqrt(0.5 * r1 * ((P.gamma - 1) * p1 + (P.gamma + 1) * p2))
    dv1s = -(2.0 * c2 / (P.gamma - 1.0)) * (1.0 - (p1 / p2) ** P.m)
    dv2r = -2.0 * (c1 + c2) * P.rgamma1
    return 1 if dvv >= dv2s else 2 if dvv >= dv1s else 3 if dvv > dv2r else 4


def x_interfaces(oracle, P, st, j=NB, k=NB):
    """The x-sweep interface states of pencil (j, k) as the reference builds
    them (order2interp on the 6-cell stencil i-3..i+2 of interface i-1/2),
    with the Riemann case and whether the high-density-peak pressure floor
    (riemannsolver.F90:279-285) raises the star pressure there.
    Returns a list of (i, state10, case, floor_raises)."""
    nx = st["p"].shape[0]
    g = np.zeros(6)
    out = []
    for i in range(NB, nx - NB + 1):
        sl = slice(i - 3, i + 3)
        pres, rho = st["p"][sl, j, k], st["rho"][sl, j, k]
        vx, vy, vz = st["vx"][sl, j, k], st["vy"][sl, j, k], st["vz"][sl, j, k]
        c = np.sqrt(P.gamma * pres / rho)
        s10, _ = oracle.order2interp(P, pres, rho, vx, vy, vz, g, c)
        res, _ = oracle.riemannsolver(P, rho, vx, s10)
        floor = False
        if rho[1] > P.dmax * rho[0] and rho[2] > P.dmax * rho[3]:
            maxp = (vx[1] - vx[2]) * max(rho[1], rho[2]) * max(0.0, vx[1] - vx[2])
            P0 = P.copy()
            P0.dmax = 1e300
            res0, _ = oracle.riemannsolver(P0, rho, vx, s10)
            floor = maxp > res0[0] and res[0] == maxp
        out.append((i, s10, riemann_case(P, s10), floor))
    return out
