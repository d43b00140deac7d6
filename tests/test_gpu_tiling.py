"""Launch shape never changes results: every tiling of the sweeps (chunks per
x wave, y/z pencil segment length, block size) gives the same bits as the
default, and boxes whose sizes sit on the x sweep's 58-cell chunk edges or
on the 3-cell minimum agree with the oracle."""
import numpy as np
import pytest

import hydropad_amd as hp
from helpers import FIELDS, NB, copy_state, rel_l1, to_hp_params
from hydropad_amd import ic

pytestmark = pytest.mark.gpu


def random_box(shape_interior, seed, amp=0.4, vamp=0.4):
    rng = np.random.default_rng(seed)
    shape = tuple(n + 2 * NB for n in shape_interior)
    st = {"p": 1.0 + amp * rng.random(shape), "rho": 1.0 + amp * rng.random(shape)}
    for k in ("vx", "vy", "vz"):
        st[k] = vamp * (rng.random(shape) - 0.5)
    st = {k: np.asfortranarray(v) for k, v in st.items()}
    return ic.periodic_fill(st)


def sweep(st, P, d, precision, tiling=None):
    with hp.Box(st["p"].shape, precision) as box:
        box.set_params(to_hp_params(P, hp.HpParams))
        if tiling is not None:
            box.set_tiling(*tiling)
        box.upload(copy_state(st))
        box.ppmsolver(d)
        return box.download()


# (x chunks per wave, y/z segment length, block size)
TILINGS = [(1, 1, 64), (3, 5, 128), (16, 7, 256), (2, 64, 192), (5, 2, 64)]


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64_exact"])
def test_tilings_are_bitwise_identical(oracle, precision):
    st = random_box((23, 17, 29), 11)
    P = oracle.params(ngrid=29, dt=0.003)
    for d in (1, 2, 3):
        base = sweep(st, P, d, precision)
        for t in TILINGS:
            out = sweep(st, P, d, precision, t)
            for k in FIELDS:
                assert np.array_equal(out[k], base[k]), (precision, d, t, k)


@pytest.mark.parametrize("shape", [(3, 3, 3), (52, 3, 4), (53, 5, 3), (3, 41, 7), (110, 3, 3)],
                         ids=lambda s: "x".join(map(str, s)))
def test_chunk_edge_and_minimum_shapes_vs_oracle(oracle, shape):
    """nx = 58 (52 + 6) makes the virtual x row exactly one chunk wide;
    53, 110 straddle chunk edges; 3 is the smallest interior the C-ABI takes."""
    st = random_box(shape, 29 + shape[0])
    P = oracle.params(ngrid=max(shape), dt=0.002)
    for d in (1, 2, 3):
        ref = copy_state(st)
        oracle.ppmsolver(P, ref, d)
        out = sweep(st, P, d, "fp64")
        for k in FIELDS:
            assert rel_l1(out[k], ref[k]) <= 1e-12, (shape, d, k, rel_l1(out[k], ref[k]))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64_exact"])
@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 5, 3), (12, 7, 9), (53, 6, 4)],
                         ids=lambda s: "x".join(map(str, s)) if isinstance(s, tuple) else s)
def test_sweeps_write_the_periodic_face_ghosts(precision, shape):
    """Every sweep writes, beside each updated cell, its periodic image in the
    face ghosts of all three axes (what exchange_mesh would copy): after one
    sweep in each direction and under two launch shapes, each face-ghost slab
    equals the interior slab it images, bit for bit.  With an interior of 3
    cells (< 2 nb) every cell has both images along that axis."""
    st = random_box(shape, 41 + sum(shape))
    N = [n for n in shape]
    sl = [slice(NB, NB + n) for n in N]
    for d in (1, 2, 3):
        for tiling in (None, (1, 1, 64)):
            with hp.Box(st["p"].shape, precision) as box:
                if tiling is not None:
                    box.set_tiling(*tiling)
                box.upload(copy_state(st))
                box.ppmsolver(d)
                out = box.download()
            for k in FIELDS:
                a = out[k]
                for ax in range(3):
                    n = N[ax]
                    face = list(sl)
                    lo_g, lo_src = list(face), list(face)
                    lo_g[ax], lo_src[ax] = slice(0, NB), slice(n, n + NB)
                    hi_g, hi_src = list(face), list(face)
                    hi_g[ax], hi_src[ax] = slice(n + NB, n + 2 * NB), slice(NB, 2 * NB)
                    assert np.array_equal(a[tuple(lo_g)], a[tuple(lo_src)]), (precision, shape, d, tiling, k, ax, "low")
                    assert np.array_equal(a[tuple(hi_g)], a[tuple(hi_src)]), (precision, shape, d, tiling, k, ax, "high")
