"""Usui tool wear on the device (mfc_tool_wear, tool_wear.cu:20-60) against the oracle's
sequential evaluation: min and max bit for bit, the mean to 1e-13 relative (the device sums in a
fixed tree order, the oracle in storage order), on a cut in contact and on set fields."""
from __future__ import annotations

import numpy as np
import pytest

import pymfc
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def check(g, o, K, alpha, tool_vel):
    a = g.tool_wear(0.0, K, alpha, tool_vel)
    b = o.tool_wear(K, alpha, tool_vel)
    assert b[3] > 0
    assert a[0] == b[0] and a[1] == b[1]
    assert a[2] == pytest.approx(b[2], rel=1e-13)
    return a, b


def test_wear_of_a_cut_in_contact():
    cfg = pymfc.cut_config(40, 12, 10.0, thermal="contact")
    g, o = pymfc.Sim(cfg), ob.OracleSim(cfg)
    g.step(200)
    o.step(200)
    v = float(np.hypot(*g.tool()["vel"]))
    check(g, o, 0.078, 2500.0, (0.0, v))
    check(g, o, 0.05, 1000.0, (v, 0.0))


def test_wear_on_set_fields_and_accumulation():
    cfg = pymfc.cut_config(12, 6, 10.0)
    g, o = pymfc.Sim(cfg), ob.OracleSim(cfg)
    N = g.N
    rng = np.random.default_rng(3)
    fields = {
        "fc": rng.normal(0.0, 1e-6, (N, 2)) * (rng.random((N, 1)) < 0.3),
        "ft": np.zeros((N, 2)),
        "vel": rng.normal(0.0, 1e-4, (N, 2)),
        "n": rng.normal(0.0, 1.0, (N, 2)),
        "T": rng.uniform(300.0, 900.0, N),
        "h": np.full(N, 1.5e-3),
    }
    fields["ft"][::7, 1] = 1e-6  # ft.y alone does not count as contact
    for f, v in fields.items():
        g.set_field(f, v)
        o.set_field(f, v)
    a, b = check(g, o, 0.078, 2500.0, (0.0, 1e-5))
    w1 = g.tool_wear(0.5, 0.078, 2500.0, (0.0, 1e-5))
    w2 = g.tool_wear(0.25, 0.078, 2500.0, (0.0, 1e-5))
    assert w1[3] == pytest.approx(a[3] + 0.5 * a[2], rel=1e-15)
    assert w2[3] == pytest.approx(w1[3] + 0.25 * w1[2], rel=1e-15)
