"""Field-by-field comparison of a HIP simulation with the oracle."""
from __future__ import annotations

import numpy as np

import pymfc

# state fields compared for every particle
STATE = ["pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "h", "flags", "id", "hash", "p", "R"]
# time derivatives: dead values for blanked particles (never read again), compared elsewhere
DERIV = ["pos_t", "vel_t", "rho_t", "S_t", "T_t"]


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.where(both_nan, 0.0, np.abs(a - b))
    scale = np.maximum(np.abs(b), 1e-300)
    r = np.where(both_nan, 0.0, d / scale)
    r = np.where((d == 0), 0.0, r)
    return float(np.nanmax(r)) if r.size else 0.0


def equal(a, b) -> bool:
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=np.asarray(a).dtype.kind == "f")


def compare(gpu, orc, fields=None) -> dict:
    """Returns {field: (bitwise_equal, max_relative_difference)}."""
    out = {}
    flags = orc.field("flags")
    live = (flags & pymfc.FLAG_BLANKED) == 0
    for f in fields or STATE + DERIV:
        a, b = gpu.field(f), orc.field(f)
        if f in DERIV:
            a, b = a[live], b[live]
        out[f] = (equal(a, b), max_rel(a, b) if a.dtype.kind == "f" else float(not equal(a, b)))
    return out


def assert_bitwise(gpu, orc, fields=None, context="") -> None:
    res = compare(gpu, orc, fields)
    bad = {f: r for f, r in res.items() if not r[0]}
    assert not bad, f"{context}: fields differ from the oracle (bitwise, max rel diff): {bad}"
