"""The reference's final-field dump format (hydropad/main.F90:125-136,
io/write_var_mpi.F90:36-65): one file per field -- rho.bin, pressure.bin,
vx.bin, vy.bin, vz.bin -- holding the interior N^3 cells as REAL(4)
(round-to-nearest from REAL(8)), Fortran order (x fastest), native byte
order, no header."""
import os

import numpy as np

FILES = (("rho", "rho.bin"), ("p", "pressure.bin"), ("vx", "vx.bin"), ("vy", "vy.bin"), ("vz", "vz.bin"))


def write_dumps(state, outdir, nbound=3):
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for key, name in FILES:
        a = state[key][nbound:-nbound, nbound:-nbound, nbound:-nbound]
        path = os.path.join(outdir, name)
        a.astype(np.float32).ravel(order="F").tofile(path)
        paths.append(path)
    return paths


def read_dumps(outdir, shape):
    """Read the five dumps back as Fortran-ordered float32 arrays of `shape`
    (the interior extents)."""
    out = {}
    for key, name in FILES:
        raw = np.fromfile(os.path.join(outdir, name), dtype=np.float32)
        out[key] = raw.reshape(shape, order="F")
    return out
