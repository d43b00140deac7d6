"""Write the frozen known-answer vectors under tests/golden/ from the
reference's OWN routines (oracle/_ref/libhydropad_ref.so, compiled from the
reference's Fortran sources by oracle/Makefile -- the reference ships no
tests or fixtures of its own).

Run where the reference tree exists (never on the GPU box):
    make -C oracle && python tests/golden/gen_golden.py

Files (all float64, numpy .npz):
  order2interp.npz  6-cell stencils -> the 10 interface states of
                    hydro/order2interp.F90 (several parameter sets)
  riemann.npz       interface states -> (pm, rhom, vxm, vym, vzm) of
                    hydro/riemannsolver.F90, all four wave patterns
  integrate.npz     fluxes + cell -> updated cell, hydro/integrate.F90, in
                    the regime e/E >= eta2 where its undefined pmean cannot
                    reach the result
  sweeps.npz        one STENCIL ppmsolver sweep per direction on small
                    periodic boxes (random, Sod, gravity, expansion terms)
  steps.npz         three full steps (ppmsolver x3 in nrot order, periodic
                    fill between sweeps, speed + timedet) of a 12^3 box
  nbody.npz         nbody/kick.F90 and nbody/drift.F90 on seeded particles
                    (inactive ones and periodic wraps included) over random
                    gforce arrays of a cubic and a non-cubic box
                    (written alone by `gen_golden.py nbody`)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import Oracle, Reference  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
NB = 3


def fields(state):
    return np.stack([state[k] for k in ("p", "rho", "vx", "vy", "vz")])


def unfields(a):
    return {k: np.asfortranarray(a[i]) for i, k in enumerate(("p", "rho", "vx", "vy", "vz"))}


def random_state(n, rng, amp=0.3, vamp=0.3):
    st = {}
    st["p"] = np.asfortranarray(1.0 + amp * rng.random((n, n, n)))
    st["rho"] = np.asfortranarray(1.0 + amp * rng.random((n, n, n)))
    for k in ("vx", "vy", "vz"):
        st[k] = np.asfortranarray(vamp * (rng.random((n, n, n)) - 0.5))
    return st


def periodic(st):
    for a in st.values():
        nx, ny, nz = a.shape
        a[:, :, :NB] = a[:, :, nz - 2 * NB:nz - NB]
        a[:, :, nz - NB:] = a[:, :, NB:2 * NB]
        a[:NB] = a[nx - 2 * NB:nx - NB]
        a[nx - NB:] = a[NB:2 * NB]
        a[:, :NB] = a[:, ny - 2 * NB:ny - NB]
        a[:, ny - NB:] = a[:, NB:2 * NB]
    return st


def param_sets(O):
    """(name, params) pairs: plain, expanding (dat != 0), gravity-agnostic."""
    P1 = O.params(ngrid=16, dt=0.02)
    P2 = O.params(ngrid=16, dt=0.004)
    P2.at, P2.dat, P2.rat = 0.05, 3.0, 1.0 / 0.05
    P2.rdtath = P2.dt / 0.06
    P3 = O.params(ngrid=48, dt=0.01)          # dx inexact through FP32 (N=48)
    return [("plain", P1), ("cosmo", P2), ("n48", P3)]


def gen_order2interp(O, R, rng):
    out = {}
    for name, P in param_sets(O):
        R.set_params(P)
        X, Y = [], []
        for i in range(200):
            kind = i % 4
            p = 1.0 + 0.3 * rng.random(6)
            rho = 1.0 + 0.3 * rng.random(6)
            v = 0.5 * (rng.random((3, 6)) - 0.5)
            g = np.zeros(6) if kind < 2 else 0.2 * (rng.random(6) - 0.5)
            if kind == 1:                       # steep jumps: flattening + steepening
                p[3:] *= 0.2
                rho[3:] *= 0.15
                v[1, :] = 0.0                   # zero transverse velocity (stale-ratio case)
            if kind == 3:                       # monotone smooth ramps
                x = np.arange(6)
                p = 1.0 + 0.05 * x + 0.01 * x ** 2
                rho = 2.0 - 0.07 * x
                v = np.stack([0.1 + 0.02 * x, -0.05 * x, 0.3 - 0.01 * x * x])
            c = np.sqrt(P.gamma * p / rho)
            inp = np.stack([p, rho, v[0], v[1], v[2], g, c])
            X.append(inp)
            Y.append(R.order2interp(*inp))
        out[name + "_in"] = np.array(X)
        out[name + "_out"] = np.array(Y)
        out[name + "_params"] = P.as_array()
    np.savez_compressed(os.path.join(OUT, "order2interp.npz"), **out)


def gen_riemann(O, R, rng):
    P = O.params(ngrid=32, dt=0.01)
    Pc = O.params(ngrid=32, dt=0.001)
    Pc.at, Pc.dat, Pc.rat = 0.1, 2.0, 10.0
    Pd = O.params(ngrid=32, dt=0.01)
    Pd.dmax = 2.0                               # exercise the high-density-peak floor
    out = {}
    for name, PP in (("plain", P), ("cosmo", Pc), ("dmax", Pd)):
        R.set_params(PP)
        X, RB, VB, Y = [], [], [], []
        tries = 0
        while len(X) < 300 and tries < 5000:
            tries += 1
            k = len(X) % 5
            pl, pr = 0.1 + 2 * rng.random(2)
            rl, rr = 0.1 + 2 * rng.random(2)
            vl, vr = rng.normal(0, 1.0, 2)
            if k == 0:                          # compression: two shocks
                vl, vr = abs(vl) + 0.5, -abs(vr) - 0.5
            elif k == 1:                        # expansion: two rarefactions
                vl, vr = -abs(vl) * 0.5, abs(vr) * 0.5
            elif k == 2:                        # vacuum
                vl, vr = -8.0 - 4 * rng.random(), 8.0 + 4 * rng.random()
            elif k == 3:                        # Sod-like, at rest
                pl, pr, rl, rr, vl, vr = 1.0, 0.1, 1.0, 0.125, 0.0, 0.0
                if rng.random() < 0.5:
                    pl, pr, rl, rr = pr, pl, rr, rl
                vl, vr = vl + rng.normal(0, 0.01), vr + rng.normal(0, 0.01)
            s = np.array([pl, pr, rl, rr, vl, vr] + list(rng.normal(0, 0.3, 4)))
            rb = 0.5 + rng.random(6)
            vb = rng.normal(0, 0.5, 6)
            if name == "dmax" and len(X) % 2 == 0:
                rb[1] = 10.0 * rb[0]            # rhobar(2) > dmax*rhobar(1)
                rb[2] = 10.0 * rb[3]            # rhobar(3) > dmax*rhobar(4)
                vb[1], vb[2] = 1.0, -1.0
            _, nit = O.riemannsolver(PP, rb, vb, s)
            if nit < 0:
                continue                        # the reference would `stop` here
            X.append(s)
            RB.append(rb)
            VB.append(vb)
            Y.append(R.riemannsolver(rb, vb, s))
        out[name + "_in"] = np.array(X)
        out[name + "_rhobar"] = np.array(RB)
        out[name + "_vxbar"] = np.array(VB)
        out[name + "_out"] = np.array(Y)
        out[name + "_params"] = PP.as_array()
    np.savez_compressed(os.path.join(OUT, "riemann.npz"), **out)


def ee1_of(P, f, st, etot, ghalf):
    """e/E after the update, following hydro/integrate.F90:147-200 (filter only)."""
    pm1, pm2, r1, r2, u1, u2, v1, v2, w1, w2 = f
    pres, rho, vx, vy, vz = st
    lam = P.rdtath
    rhonew = rho + lam * (r1 * u1 - r2 * u2) * P.rdx
    if rhonew <= 0:
        return -1.0
    grav = (rho + rhonew) * ghalf * 0.5
    vxnew = (vx * rho + lam * ((r1 * u1 * u1 + pm1) - (r2 * u2 * u2 + pm2)) * P.rdx + lam * grav) / rhonew
    vynew = (rho * vy + lam * (v1 * u1 * r1 - v2 * u2 * r2) * P.rdx) / rhonew
    vznew = (rho * vz + lam * (w1 * u1 * r1 - w2 * u2 * r2) * P.rdx) / rhonew
    fl1 = (P.gamma * pm1 * P.rgamma1 + 0.5 * r1 * (u1 * u1 + v1 * v1 + w1 * w1)) * u1
    fl2 = (P.gamma * pm2 * P.rgamma1 + 0.5 * r2 * (u2 * u2 + v2 * v2 + w2 * w2)) * u2
    grav = (rho * vx + rhonew * vxnew) * ghalf * 0.5
    etotnew = etot + lam * (fl1 - fl2) * P.rdx + lam * grav
    eeaux = etotnew - 0.5 * rhonew * (vxnew * vxnew + vynew * vynew + vznew * vznew)
    return eeaux / etotnew if etotnew > 0 else -1.0


def gen_integrate(O, R, rng):
    P = O.params(ngrid=64, dt=0.005)
    R.set_params(P)
    F, S, E, Y = [], [], [], []
    while len(F) < 600:
        f = np.empty(10)
        f[0:2] = 0.2 + rng.random(2)            # pm1, pm2
        f[2:4] = 0.2 + rng.random(2)            # rhom1, rhom2
        f[4:10] = rng.normal(0, 0.4, 6)         # vxm, vym, vzm (1,2)
        st = np.array([0.2 + rng.random(), 0.2 + rng.random()] + list(rng.normal(0, 0.5, 3)))
        ghalf = 0.0 if len(F) % 3 else rng.normal(0, 0.3)
        etot = O.total(P, *st)
        eint = O.internal(P, st[0])
        if ee1_of(P, f, st, etot, ghalf) < 0.2:
            continue                            # keep e/E well above eta2 (pmean-free regime)
        F.append(f)
        S.append(st)
        E.append([etot, eint, ghalf])
        Y.append(R.integrate(f, st, etot, eint, ghalf))
    np.savez_compressed(os.path.join(OUT, "integrate.npz"), flux=np.array(F), state=np.array(S),
                        e=np.array(E), out=np.array(Y), params=P.as_array())


def gen_sweeps(O, R, rng):
    out = {}
    n = 8 + 2 * NB
    cases = []
    cases.append(("rnd", periodic(random_state(n, rng)), None, O.params(ngrid=8, dt=0.01)))
    sod = {k: np.zeros((n, n, n), order="F") for k in ("p", "rho", "vx", "vy", "vz")}
    x = (np.arange(n) - NB + 0.5) / 8
    sod["rho"][...] = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
    sod["p"][...] = np.where(x < 0.5, 1.0, 0.1)[:, None, None]
    cases.append(("sod", sod, None, O.params(ngrid=8, dt=0.03)))
    g = np.asfortranarray(0.05 * (rng.random((3, n, n, n)) - 0.5))
    cases.append(("grav", periodic(random_state(n, rng)), g, O.params(ngrid=8, dt=0.01)))
    Pc = O.params(ngrid=8, dt=0.002)
    Pc.at, Pc.dat, Pc.rat, Pc.rdtath = 0.08, 2.5, 1 / 0.08, 0.002 / 0.085
    cases.append(("cosmo", periodic(random_state(n, rng)), None, Pc))
    for name, st, gf, P in cases:
        R.set_params(P)
        out[name + "_in"] = fields(st)
        out[name + "_params"] = P.as_array()
        if gf is not None:
            out[name + "_gforce"] = gf
        for d in (1, 2, 3):
            a = {k: v.copy(order="F") for k, v in st.items()}
            R.ppmsolver(a, d, gforce=gf)
            out[f"{name}_out{d}"] = fields(a)[:, NB:-NB, NB:-NB, NB:-NB]
    np.savez_compressed(os.path.join(OUT, "sweeps.npz"), **out)


def gen_steps(O, R, rng):
    """Three steps of the time loop with the reference's sweeps; the periodic
    fill, speed and timedet are the few lines of exchange_mesh.F90:47-123,
    speed.F90:22-31 and functions.F90:15-32."""
    n = 12 + 2 * NB
    st = periodic(random_state(n, rng, amp=0.5, vamp=0.6))
    st0 = {k: v.copy(order="F") for k, v in st.items()}
    P = O.params(ngrid=12)
    cour = 0.8
    nb = NB

    def speed(s):
        sl = (slice(nb, -nb),) * 3
        vs = np.sqrt(P.gamma * s["p"][sl] / s["rho"][sl])
        return float(np.max(np.maximum(np.maximum(np.abs(s["vx"][sl]) + vs, np.abs(s["vy"][sl]) + vs),
                                       np.abs(s["vz"][sl]) + vs)))

    t = 0.0
    vm = speed(st)
    dt = 1.0 / (vm / (cour * P.dx))
    t = t + dt
    P.dt, P.rdtath = dt, dt
    dts = [dt]
    order = {1: (1, 2, 3), 2: (3, 2, 1), 3: (2, 3, 1), 4: (1, 3, 2), 5: (3, 1, 2), 6: (2, 1, 3)}
    for nstep in (1, 2, 3):
        nrot = 1 if nstep == 1 else nstep % 6 + 1
        R.set_params(P)
        for d in order[nrot]:
            R.ppmsolver(st, d)
            periodic(st)
        vm = speed(st)
        dt = 1.0 / (vm / (cour * P.dx))
        t = t + dt
        P.dt, P.rdtath = dt, dt
        dts.append(dt)
    np.savez_compressed(os.path.join(OUT, "steps.npz"), state_in=fields(st0), state_out=fields(st),
                        dts=np.array(dts), t=np.array([t]), cour=np.array([cour]))


def nbody_case(rng, ngrid, npart):
    """Random gforce (ghosts included), particles inside [0, boxsize) plus a
    few at the box edges and inactive ones (x = -1)."""
    shape = tuple(n + 2 * NB for n in ngrid)
    g = np.asfortranarray(rng.standard_normal((3,) + shape))
    bs = np.array([np.float32(n) / np.float32(max(ngrid)) for n in ngrid], dtype=np.float64)
    x = rng.random((npart, 3)) * bs
    x[0] = 0.0
    x[1] = np.nextafter(bs, 0.0)
    x[2] = 0.5 * bs
    x[3:6, 0] = -1.0
    v = rng.standard_normal((npart, 3))
    v[6] = -2.0 * bs                       # drifts out below 0 and wraps
    v[7] = 2.0 * bs                        # and above boxsize
    return g, np.ascontiguousarray(x), np.ascontiguousarray(v)


def gen_nbody(O, R, rng):
    out = {}
    for name, ngrid in (("cubic", (8, 8, 8)), ("noncubic", (6, 8, 5))):
        g, x, v = nbody_case(rng, ngrid, 120)
        dat, dtat, rdtath = 0.35, 0.02, 0.05
        vk = v.copy()
        R.kick(x, vk, g, dat, dtat)
        xd = x.copy()
        R.drift(xd, v * 0.0 + v, rdtath, ngrid)
        out.update({name + "_gforce": g, name + "_ppos": x, name + "_pvel": v, name + "_kick": vk,
                    name + "_drift": xd, name + "_scal": np.array([dat, dtat, rdtath]),
                    name + "_ngrid": np.array(ngrid)})
    np.savez_compressed(os.path.join(OUT, "nbody.npz"), **out)


def main():
    O = Oracle()
    if not Reference.available():
        raise SystemExit("oracle/_ref not built (needs the reference tree): make -C oracle ref")
    R = Reference()
    if sys.argv[1:] == ["nbody"]:
        gen_nbody(O, R, np.random.default_rng(20261203))
        return
    rng = np.random.default_rng(20261202)
    gen_order2interp(O, R, rng)
    gen_riemann(O, R, rng)
    gen_integrate(O, R, rng)
    gen_sweeps(O, R, rng)
    gen_steps(O, R, rng)
    gen_nbody(O, R, np.random.default_rng(20261203))
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
