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
