"""mfc_tool_wear_noalloc (the entry point that allocates nothing) against mfc_tool_wear through
the C ABI: the same kernels, so bit-identical results and accumulation, on two calls in a row
with the second handle smaller than the first, after a larger handle filled the shared scratch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pymfc

pytestmark = pytest.mark.gpu


def wear_pair(cfg, steps, dt):
    """(out of mfc_tool_wear, out of mfc_tool_wear_noalloc) on two identical runs"""
    a, b = pymfc.Sim(cfg), pymfc.Sim(cfg)
    a.step(steps)
    b.step(steps)
    tv = np.array([0.0, float(np.hypot(*a.tool()["vel"]))])
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    outs = []
    for s, fn in ((a, pymfc.lib().mfc_tool_wear), (b, pymfc.lib().mfc_tool_wear_noalloc)):
        o = np.zeros(4)
        for _ in range(2):  # the second call accumulates dt * mean on the handle
            assert fn(s._h, 0.078, 2500.0, p(tv), dt, p(o)) == 0, pymfc.lib().mfc_last_error()
        outs.append(o)
    n = a.N
    a.close()
    b.close()
    return outs[0], outs[1], n


def test_noalloc_wear_equals_wear_on_two_calls():
    big = pymfc.cut_config(120, 40, 10.0, thermal="contact", speed_m_min=1000.0)
    small = pymfc.cut_config(40, 12, 10.0, thermal="contact", speed_m_min=1000.0)
    w1, n1, N1 = wear_pair(big, 200, 1e-3)
    assert w1[3] > 0 and np.array_equal(w1, n1), (w1, n1)
    w2, n2, N2 = wear_pair(small, 150, 2e-3)
    assert N2 < N1 // 2
    assert w2[3] > 0 and np.array_equal(w2, n2), (w2, n2)


def test_noalloc_wear_refuses_handles_beyond_its_scratch():
    cfg = pymfc.cut_config(2050, 1030, 1.0, thermal="adiabatic")
    s = pymfc.Sim(cfg)
    assert s.N > 256 * 8192
    o = np.zeros(4)
    tv = np.zeros(2)
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    rc = pymfc.lib().mfc_tool_wear_noalloc(s._h, 0.078, 2500.0, p(tv), 0.0, p(o))
    assert rc == -4  # MFC_E_RANGE
    assert b"at most 8192 tiles" in pymfc.lib().mfc_last_error()
    s.close()
