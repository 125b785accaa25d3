"""The oracle's Usui tool-wear evaluation (tool_wear.cu:20-60), pinned by special cases: which
particles count as in contact (the reference tests fc.y twice and never ft.y), the component-wise
product in the relative velocity v - v*n (a tangential projection would give another value), the
pressure |fc| / (2 pi h), the temperature factor exp(alpha / T), min / max / mean over the contact
set and linearity in K. Expected values are evaluated with numpy's libm."""
from __future__ import annotations

import math

import numpy as np
import pytest

import pymfc
import oracle_binding as ob


def small():
    o = ob.OracleSim(pymfc.cut_config(12, 6, 10.0))
    N = o.N
    zero2 = np.zeros((N, 2))
    for f in ("fc", "ft", "vel", "n"):
        o.set_field(f, zero2)
    o.set_field("T", np.full(N, 300.0))
    o.set_field("h", np.full(N, 1.5e-3))
    return o


def rate(K, alpha, T, h, fc, v, n, tool_vel=(0.0, 0.0)):
    vx, vy = v[0] - tool_vel[0], v[1] - tool_vel[1]
    rx, ry = vx - vx * n[0], vy - vy * n[1]
    p = math.hypot(fc[0], fc[1]) / (2 * h * math.pi)
    return K * math.exp(alpha / T) * p * math.hypot(rx, ry)


def put(o, name, i, value):
    a = o.field(name)
    a[i] = value
    o.set_field(name, a)


def test_no_contact_gives_zeros():
    o = small()
    assert np.array_equal(o.tool_wear(), np.zeros(4))


def test_contact_test_ignores_ft_y():
    o = small()
    put(o, "ft", 3, [0.0, 1e-6])
    assert o.tool_wear()[3] == 0  # ft.y alone: not in contact (tool_wear.cu:36)
    put(o, "ft", 3, [1e-6, 0.0])
    out = o.tool_wear()
    assert out[3] == 1 and out[0] == 0.0 and out[1] == 0.0  # in contact, but |fc| = 0: zero rate


def test_rate_of_one_particle():
    """v = (3, 4) 1e-4, n = (0.6, 0.8): the component-wise v - v*n is (1.2, 0.8) 1e-4, while the
    tangential projection v - (v.n) n would be zero."""
    o = small()
    fc, v, n, T, h = [3e-6, 4e-6], [3e-4, 4e-4], [0.6, 0.8], 650.0, 2e-3
    put(o, "fc", 5, fc)
    put(o, "vel", 5, v)
    put(o, "n", 5, n)
    put(o, "T", 5, T)
    put(o, "h", 5, h)
    K, alpha = 0.078, 2500.0
    out = o.tool_wear(K, alpha)
    want = rate(K, alpha, T, h, fc, v, n)
    assert want > 0
    assert out[3] == 1
    for q in range(3):
        assert out[q] == pytest.approx(want, rel=1e-14)


def test_min_max_mean_over_the_contact_set():
    o = small()
    rows = [(2, [1e-6, 0.0], [1e-4, 0.0], [0.0, 1.0], 400.0),
            (7, [0.0, 2e-6], [0.0, -2e-4], [1.0, 0.0], 500.0),
            (11, [3e-6, 3e-6], [5e-5, 5e-5], [0.5, 0.5], 350.0)]
    tool_vel = (1e-5, -2e-5)
    for i, fc, v, n, T in rows:
        put(o, "fc", i, fc)
        put(o, "vel", i, v)
        put(o, "n", i, n)
        put(o, "T", i, T)
    K, alpha = 0.05, 1800.0
    out = o.tool_wear(K, alpha, tool_vel)
    w = [rate(K, alpha, T, 1.5e-3, fc, v, n, tool_vel) for _, fc, v, n, T in rows]
    assert out[3] == 3
    assert out[0] == pytest.approx(min(w), rel=1e-14)
    assert out[1] == pytest.approx(max(w), rel=1e-14)
    assert out[2] == pytest.approx(sum(w) / 3, rel=1e-14)
    out2 = o.tool_wear(2 * K, alpha, tool_vel)  # linear in K
    np.testing.assert_allclose(out2[:3], 2 * out[:3], rtol=1e-15)
