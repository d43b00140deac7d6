"""The Fortran drop-in (hydropad_amd/fortran): the ISO_C_BINDING module binds
only entry points the C header declares, the reference-format dumps
(hydropad/main.F90:125-136, io/write_var_mpi.F90:55-65) round-trip, and on
the GPU the evolve_sys-shaped Fortran driver reproduces the Python host path
bit for bit and the oracle to FP32 rounding."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import FIELDS, NB, copy_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "hydropad_amd", "fortran")
DRIVER = os.path.join(FDIR, "hydropad_driver")


def test_fortran_bindings_match_header():
    src = open(os.path.join(FDIR, "hydropad_gpu.F90")).read()
    hdr = open(os.path.join(ROOT, "include", "hydropad_gpu.h")).read()
    names = re.findall(r'bind\(C,\s*name="(\w+)"\)', src)
    assert len(names) >= 14
    for n in names:
        assert re.search(r"\b" + n + r"\s*\(", hdr), n
    # hp_params field order of the bind(C) type = the C struct's
    ftype = re.search(r"type, bind\(C\), public :: hp_params_t(.*?)end type", src, re.S).group(1)
    ffields = re.findall(r"(\w+)", " ".join(l.split("::")[1] for l in ftype.strip().splitlines()))
    cstruct = re.search(r"typedef struct hp_params\s*\{(.*?)\}\s*hp_params;", hdr, re.S).group(1)
    cfields = re.findall(r"(\w+)", cstruct.replace("double", ""))
    assert ffields == cfields


def test_dump_format_roundtrip(tmp_path):
    from hydropad_amd import dumps
    N = 5
    n = N + 2 * NB
    rng = np.random.default_rng(3)
    st = {k: np.asfortranarray(rng.standard_normal((n, n, n))) for k in FIELDS}
    paths = dumps.write_dumps(st, str(tmp_path))
    assert [os.path.basename(p) for p in paths] == ["rho.bin", "pressure.bin", "vx.bin", "vy.bin", "vz.bin"]
    for p in paths:
        assert os.path.getsize(p) == 4 * N ** 3          # REAL(4), interior only, no header
    raw = np.fromfile(os.path.join(str(tmp_path), "rho.bin"), dtype=np.float32)
    # Fortran order: x fastest
    assert raw[0] == np.float32(st["rho"][NB, NB, NB])
    assert raw[1] == np.float32(st["rho"][NB + 1, NB, NB])
    assert raw[N] == np.float32(st["rho"][NB, NB + 1, NB])
    back = dumps.read_dumps(str(tmp_path), (N, N, N))
    for k in FIELDS:
        assert np.array_equal(back[k], st[k][NB:-NB, NB:-NB, NB:-NB].astype(np.float32))


def _run_driver(n, steps, prec, outdir, mode="resident"):
    r = subprocess.run([DRIVER, str(n), str(steps), prec, outdir, mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"t=\s*(\S+)\s+dt=\s*(\S+)", r.stdout)
    return float(m.group(1)), float(m.group(2))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "exact"])
def test_fortran_driver_matches_host_path_and_oracle(tmp_path, prec, oracle):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no flang)")
    import hydropad_amd as hp
    from hydropad_amd import dumps, ic
    N, K = 32, 12
    print(f"fortran driver {prec} N={N} steps={K}", flush=True)
    t_f, dt_f = _run_driver(N, K, prec, str(tmp_path))
    got = dumps.read_dumps(str(tmp_path), (N, N, N))

    st = ic.sod(N)
    with hp.Box(N, "fp64" if prec == "fp64" else "fp64_exact") as box:
        box.upload(copy_state(st))
        dt0, t0 = box.timedet(box.speed(), 0.8)
        t_h, dt_h = box.evolve(1, K, 0.8, t0, dt0)
        host = box.download()
    assert t_f == t_h and dt_f == dt_h
    for k in FIELDS:
        assert np.array_equal(got[k], host[k][NB:-NB, NB:-NB, NB:-NB].astype(np.float32)), k

    P = oracle.params(ngrid=N)
    ref = copy_state(st)
    dt_o, t_o = oracle.timedet(P, oracle.speed(P, ref), 0.8, 1e30, 0.0)
    oracle.evolve(P, ref, 1, K, 0.8, 1e30, t_o)
    for k in FIELDS:
        a = got[k].astype(np.float64)
        b = ref[k][NB:-NB, NB:-NB, NB:-NB]
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 2 ** -23 * scale * 2, k


@pytest.mark.gpu
def test_fortran_per_sweep_dropin_matches_resident(tmp_path):
    """ppmsolver_gpu with host arrays at every ppmsolver call site of
    ppm_lxlylz (upload, sweep, download each time, host speed) gives the same
    dumps and times as the device-resident loop."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no flang)")
    from hydropad_amd import dumps
    a, b = tmp_path / "res", tmp_path / "swp"
    a.mkdir()
    b.mkdir()
    print("fortran driver resident vs sweeps, N=24, 9 steps", flush=True)
    ta = _run_driver(24, 9, "fp64", str(a), "resident")
    tb = _run_driver(24, 9, "fp64", str(b), "sweeps")
    assert ta == tb
    da, db = dumps.read_dumps(str(a), (24, 24, 24)), dumps.read_dumps(str(b), (24, 24, 24))
    for k in FIELDS:
        assert np.array_equal(da[k], db[k]), k


@pytest.mark.gpu
def test_fortran_nbody_dropins_match_oracle(tmp_path, oracle):
    """kick_gpu / drift_gpu with the reference's argument lists (nbody/kick.F90:1,
    nbody/drift.F90:1) and the device-resident nbody_step_gpu, three steps of
    kick, drift, kick under the Gaussian gravity set-up, against the oracle
    (pinned bit for bit to the reference's kick and drift): bit for bit."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no flang)")
    from hydropad_amd import ic
    N, K, NP = 16, 3, 3000
    print(f"fortran nbody drop-ins N={N} steps={K} np={NP}", flush=True)
    shape = (N + 2 * NB,) * 3
    g = ic.gaussian_gforce(shape)
    x, v = ic.particles((N, N, N), NP, 21, vamp=1.0)
    x[:2, 0] = -1.0                                         # inactive particles
    g.ravel(order="F").tofile(str(tmp_path / "gforce.bin"))
    x.T.ravel(order="F").tofile(str(tmp_path / "ppos.bin"))  # (3, np) Fortran order
    v.T.ravel(order="F").tofile(str(tmp_path / "pvel.bin"))
    r = subprocess.run([DRIVER, str(N), str(K), "fp64", str(tmp_path), "nbody", "0", str(NP)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "particles_outside=0" in r.stdout
    P = oracle.params(ngrid=N)
    P.dat, P.rdtath = 0.2, 0.04
    xo, vo = x.copy(), v.copy()
    for _ in range(K):
        assert oracle.kick(P, xo, vo, g, 0.015) == 0
        oracle.drift(P, xo, vo, (N, N, N))
        assert oracle.kick(P, xo, vo, g, 0.015) == 0
    for form in ("call", "res"):
        xf = np.fromfile(str(tmp_path / f"ppos_{form}.bin")).reshape(NP, 3)
        vf = np.fromfile(str(tmp_path / f"pvel_{form}.bin")).reshape(NP, 3)
        assert np.array_equal(xf, xo), form
        assert np.array_equal(vf, vo), form
