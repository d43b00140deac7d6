"""The general-gamma kernels (FastMath: the Riemann solver's powers through the
range-reduced ln/exp pair instead of gamma = 5/3's roots) against the oracle
at the same gamma.  gamma = 1.4 (air), and 1.66666667 -- the reference's own
namelist value (input.nml:39), which is not 5/3 and so never takes the
constant-folded kernels.  Single box in the three forms, the comoving loop,
and a batch of boxes with either gamma: FP64 forms at relative L1 <= 1e-12
per field, mixed at L1(rho) <= 1e-5."""
import numpy as np
import pytest

import hydropad_amd as hp
from helpers import FIELDS, copy_state, rel_l1, rel_linf
from hydropad_amd import batch, ic

pytestmark = pytest.mark.gpu

GAMMAS = [1.4, 1.66666667]


def oracle_evolve(oracle, st, N, K, gamma, cour=0.8):
    ref = copy_state(st)
    P = oracle.params(gamma=gamma, ngrid=N)
    dt, t = oracle.timedet(P, oracle.speed(P, ref), cour, 1e30, 0.0)
    oracle.evolve(P, ref, 1, K, cour, 1e30, t)
    return ref, P


@pytest.mark.parametrize("gamma", GAMMAS)
@pytest.mark.parametrize("precision", ["fp64", "fp64_exact", "mixed"])
def test_general_gamma_evolve_vs_oracle(oracle, gamma, precision):
    print(f"gamma={gamma} {precision}", flush=True)
    N, K = 24, 20
    st = ic.sod(N)
    rng = np.random.default_rng(int(gamma * 1e8) % 1000)
    st["vy"] = np.asfortranarray(0.05 * rng.standard_normal(st["vy"].shape))   # not planar
    ref, P = oracle_evolve(oracle, st, N, K, gamma)
    with hp.Box(N, precision, gamma=gamma) as box:
        assert box.params.gamma == P.gamma and box.params.m == P.m and box.params.gf == P.gf
        box.upload(copy_state(st))
        dt, t = box.timedet(box.speed(), 0.8)
        box.evolve(1, K, 0.8, t, dt)
        out = box.download()
        c = box.counters()
    assert c.newton_nonconv == 0
    if precision == "mixed":
        assert rel_l1(out["rho"], ref["rho"]) <= 1e-5
    else:
        for k in FIELDS:
            assert rel_l1(out[k], ref[k]) <= 1e-12, (k, rel_l1(out[k], ref[k]), rel_linf(out[k], ref[k]))


def test_general_gamma_comoving_vs_oracle(oracle):
    gamma, N, K = 1.4, 20, 10
    st = ic.grf_box(N, 31, gamma=gamma)
    ca, tstop = ic.eds_schedule(1.0 / 51.0, 1.0 / 11.0, 200)
    P = oracle.params(gamma=gamma, ngrid=N)
    s_ref = oracle.cosmo_init(1.0 / 51.0, 0.8, ca, tstop)
    ref = copy_state(st)
    assert oracle.evolve_cosmo(P, ref, 1, K, s_ref) == 0
    s = hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop)
    with hp.Box(N, "fp64", gamma=gamma) as box:
        box.upload(copy_state(st))
        box.evolve_cosmo(1, K, s)
        out = box.download()
    assert s.nstep == s_ref.nstep == K
    for k in FIELDS:
        assert rel_l1(out[k], ref[k]) <= 1e-12, (k, rel_l1(out[k], ref[k]))


def test_batch_with_mixed_gammas_vs_oracle(oracle):
    """One batch holding gamma = 5/3 and gamma = 1.4 boxes: each box's own
    scalars select the constant-folded or the general kernels."""
    cases = [(12, 1, 8, 5.0 / 3.0), (16, 2, 10, 1.4), (14, 3, 6, 1.4), (10, 4, 9, 5.0 / 3.0)]
    states = [ic.random_smooth(N, seed) for N, seed, _, _ in cases]
    with batch.Batch(max_boxes=8) as b:
        ids = [b.add(copy_state(st), nsteps=K, gamma=g) for st, (_, _, K, g) in zip(states, cases)]
        b.run_until_done()
        outs = [b.download(i) for i in ids]
    for (N, seed, K, g), s0, out in zip(cases, states, outs):
        print(f"batch gamma={g:.4f} N={N} steps={K}", flush=True)
        ref, _ = oracle_evolve(oracle, s0, N, K, g)
        for k in FIELDS:
            assert rel_l1(out[k], ref[k]) <= 1e-12, (g, N, k, rel_l1(out[k], ref[k]))
