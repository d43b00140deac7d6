"""The batched ensemble and the serving loop on the GPU.

A box advanced inside a batch runs the same per-cell arithmetic as the single
box path (HP_FP64 / HP_MIXED), whatever else is in the batch, whatever the
tiling, eager or replayed from a hipGraph: these tests hold the batch to the
oracle (FP64: rel L1 <= 1e-12; mixed: L1(rho) <= 1e-5), to the single-box
path, and to itself under re-tiling, graph replay and continuous batching
(bit for bit)."""
import numpy as np
import pytest

import hydropad_amd as hp
from helpers import FIELDS, NB, copy_state, rel_l1, rel_linf
from hydropad_amd import batch, ic

pytestmark = pytest.mark.gpu


def single_box(st, nsteps, precision):
    N = st["p"].shape[0] - 2 * NB
    with hp.Box(N, precision) as box:
        box.upload(copy_state(st))
        dt, t = box.timedet(box.speed(), 0.8)
        box.evolve(1, nsteps, 0.8, t, dt)
        return box.download()


def same(a, b):
    return all(np.array_equal(a[k][NB:-NB, NB:-NB, NB:-NB], b[k][NB:-NB, NB:-NB, NB:-NB]) for k in FIELDS)


def test_batch_matches_oracle_and_single_boxes(oracle):
    cases = [(12, 1, 7), (16, 2, 12), (20, 3, 9), (14, 4, 10)]
    states = [ic.random_smooth(N, seed) for N, seed, _ in cases]
    with batch.Batch(max_boxes=8) as b:
        ids = [b.add(copy_state(st), nsteps=K) for st, (_, _, K) in zip(states, cases)]
        st = b.run_until_done()
        outs = [b.download(i) for i in ids]
    for (N, seed, K), s0, out, i in zip(cases, states, outs, ids):
        print(f"batch box N={N} steps={K}", flush=True)
        assert st[i].steps_done == K and st[i].newton_nonconv == 0
        ref = copy_state(s0)
        P = oracle.params(ngrid=N)
        dt, t = oracle.timedet(P, oracle.speed(P, ref), 0.8, 1e30, 0.0)
        oracle.evolve(P, ref, 1, K, 0.8, 1e30, t)
        for k in FIELDS:
            assert rel_l1(out[k], ref[k]) <= 1e-12, (N, k, rel_l1(out[k], ref[k]), rel_linf(out[k], ref[k]))
        single = single_box(s0, K, "fp64")            # built separately: FMA contraction may differ
        for k in FIELDS:
            assert rel_l1(out[k], single[k]) <= 1e-12, (N, k)


def test_batch_mixed_precision_and_comoving(oracle):
    ca, tstop = ic.eds_schedule(1.0 / 51.0, 1.0 / 11.0, 200)
    g = ic.grf_box(16, 8)
    r = ic.random_smooth(12, 9)
    with batch.Batch(max_boxes=4) as b:
        s = hp.cosmo_init(1.0 / 51.0, 0.8, ca, tstop)
        i_c = b.add(copy_state(g), nsteps=8, cosmo=s)
        i_m = b.add(copy_state(r), nsteps=6, precision="mixed")
        i_cm = b.add(copy_state(g), nsteps=8, cosmo=s, precision="mixed")
        st = b.run_until_done()
        out_c, out_m, out_cm = b.download(i_c), b.download(i_m), b.download(i_cm)
    # comoving FP64 against the oracle's evolve_sys with COSMO
    ref = copy_state(g)
    so = oracle.cosmo_init(1.0 / 51.0, 0.8, ca, tstop)
    oracle.evolve_cosmo(oracle.params(ngrid=16), ref, 1, 8, so)
    assert st[i_c].t == pytest.approx(so.t, rel=1e-14)
    for k in FIELDS:
        assert rel_l1(out_c[k], ref[k]) <= 1e-12, k
    assert rel_l1(out_cm["rho"], ref["rho"]) <= 1e-5
    # static mixed against the oracle
    ref = copy_state(r)
    P = oracle.params(ngrid=12)
    dt, t = oracle.timedet(P, oracle.speed(P, ref), 0.8, 1e30, 0.0)
    oracle.evolve(P, ref, 1, 6, 0.8, 1e30, t)
    assert rel_l1(out_m["rho"], ref["rho"]) <= 1e-5


def test_batch_tiling_and_graph_invariance():
    """Re-tiling and hipGraph replay change nothing, bit for bit; every replay
    is checked against an eager batch, with eager work in between."""
    sts = [ic.random_smooth(N, s) for N, s in ((18, 1), (10, 2), (24, 3))]
    with batch.Batch(max_boxes=4) as a, batch.Batch(max_boxes=4) as g:
        ia = [a.add(copy_state(s), nsteps=20) for s in sts]
        ig = [g.add(copy_state(s), nsteps=20) for s in sts]
        a.set_tuning(x_chunks=1, yz_segments=1, waves_per_cu=4)
        g.set_tuning(x_chunks=16, yz_segments=4, waves_per_cu=16)
        g.use_graph(True)
        for rep in range(5):
            print(f"graph replay {rep}", flush=True)
            a.step(1)
            g.step(1)
            a.poll()
            g.poll()
            for x, y in zip(ia, ig):
                assert same(a.download(x), g.download(y)), (rep, x)
        g.step(3)
        a.step(3)
        for x, y in zip(ia, ig):
            assert same(a.download(x), g.download(y))


def test_autotune_keeps_results():
    sts = [ic.random_smooth(N, s) for N, s in ((16, 4), (20, 5))]
    with batch.Batch(max_boxes=4) as a, batch.Batch(max_boxes=4) as t:
        ia = [a.add(copy_state(s), nsteps=40) for s in sts]
        it = [t.add(copy_state(s), nsteps=40) for s in sts]
        best, ms, n = t.autotune(steps_per_trial=1)
        assert n == 22 and ms > 0.0
        assert best["x_chunks"] in (1, 2, 4, 8, 16) and best["waves_per_cu"] in (4, 8, 16)
        a.step(n)
        for x, y in zip(ia, it):
            assert same(a.download(x), t.download(y))


def test_device_sod_matches_host_generator():
    for seed in (3, 17, 2024):
        axis, x0, flip = batch.sod_of_seed(seed)
        left, right = ((1.0, 1.0), (0.125, 0.1))
        if flip:
            left, right = right, left
        host = ic.sod(20, axis=axis, left=left, right=right, x0=x0)
        with batch.Batch(max_boxes=2) as b:
            i = b.add_sod(20, seed, nsteps=1)
            dev = b.download(i)
        for k in FIELDS:
            assert np.array_equal(dev[k], host[k]), (seed, k)


def test_serve_result_same_alone_and_in_mix(oracle):
    """Continuous batching: a request's result does not depend on what else
    is resident -- identical bit for bit alone and in a Poisson mix of
    requests of other sizes, step counts and precisions."""
    rng = np.random.default_rng(7)
    probe = {"n": 24, "nsteps": 30, "precision": "mixed", "seed": 99, "arrival_s": 0.0}
    alone, st1, f1 = batch.serve([probe], keep_fields=(0,))
    assert alone[0]["rc"] == 0 and alone[0]["steps"] == 30
    mix = []
    t = 0.0
    for i in range(12):
        t += rng.exponential(0.004)
        mix.append({"n": int(rng.choice([12, 16, 20, 28, 32])), "nsteps": int(rng.integers(5, 40)),
                    "precision": "fp64" if i % 2 else "mixed", "seed": 1000 + i, "arrival_s": t})
    mix.insert(5, dict(probe, arrival_s=mix[4]["arrival_s"]))
    res, stats, f2 = batch.serve(mix, keep_fields=(5,), hbm_budget=64 << 20, max_boxes=6)
    assert all(r["rc"] == 0 for r in res)
    assert stats["completed"] == len(mix) and stats["max_concurrent"] >= 2
    assert stats["peak_bytes"] <= 64 << 20
    assert 0.0 < stats["p50_s"] <= stats["p99_s"]
    for k in FIELDS:
        assert np.array_equal(f1[0][k], f2[5][k]), k
    # and the request alone agrees with the single-box mixed path to FP32
    # rounding accumulated over 30 steps (the two compile the FP32
    # reconstruction and Riemann sampling separately, so FMA contraction may
    # differ and a last-bit difference can round an FP32 result the other
    # way) and with the FP64 oracle to the mixed bar
    axis, x0, flip = batch.sod_of_seed(99)
    left, right = ((1.0, 1.0), (0.125, 0.1))
    if flip:
        left, right = right, left
    s0 = ic.sod(24, axis=axis, left=left, right=right, x0=x0)
    single = single_box(s0, 30, "mixed")
    for k in FIELDS:
        assert rel_l1(f1[0][k], single[k]) <= 1e-6, k
    ref = copy_state(s0)
    P = oracle.params(ngrid=24)
    dt, t = oracle.timedet(P, oracle.speed(P, ref), 0.8, 1e30, 0.0)
    oracle.evolve(P, ref, 1, 30, 0.8, 1e30, t)
    assert rel_l1(f1[0]["rho"], ref["rho"]) <= 1e-5


def test_serve_admission_by_budget():
    """With room for one 32^3 box at a time, requests queue and run one by one."""
    one = batch.box_bytes((38, 38, 38))
    reqs = [{"n": 32, "nsteps": 4, "precision": "fp64", "seed": i, "arrival_s": 0.0} for i in range(4)]
    res, stats, _ = batch.serve(reqs, hbm_budget=one + one // 2, max_boxes=8)
    assert stats["max_concurrent"] == 1 and stats["completed"] == 4
    admits = sorted(r["admit_s"] for r in res)
    dones = sorted(r["done_s"] for r in res)
    for a, d in zip(admits[1:], dones[:-1]):
        assert a >= d
    big = [{"n": 64, "nsteps": 2, "precision": "fp64", "seed": 1, "arrival_s": 0.0}]
    res, stats, _ = batch.serve(big, hbm_budget=one, max_boxes=2)
    assert res[0]["rc"] == hp._capi.HP_E_NOMEM and stats["completed"] == 0
