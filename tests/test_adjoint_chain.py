"""The whole reverse-mode chain restated on the host from the blocks of mfc_adjoint.h
(tests/harness/adj_chain.cpp): its forward must be the oracle's step, and its gradient must
match central differences of the oracle. The GPU reverse mode is then checked against this
host chain (tests/test_gpu_adjoint.py)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_binding as ob
import pymfc
from test_adjoint_blocks import L

L.adjh_chain.argtypes = [C.POINTER(pymfc.Config), C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5
L.adjh_chain_sub.argtypes = [C.POINTER(pymfc.Config), C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4

THETA = pymfc.THETA


def adj_config(thermal="adiabatic"):
    cfg = pymfc.cut_config(40, 12, 10.0, thermal=thermal, speed_m_min=1000.0)
    cfg.promote_fp64 = 1
    cfg.secant_tol = 1e-14
    return cfg


def theta_of(cfg):
    return np.array([getattr(cfg, n) for n in THETA])


def with_theta(cfg, th):
    c = cfg.copy()
    for n, v in zip(THETA, th):
        setattr(c, n, float(v))
    return c


def host_chain(cfg, nsteps, w0, w1, cot=None):
    """(Fmean, grad or None, final sorted state, force history)"""
    N = pymfc.scenario(cfg)["info"].n_particles
    Fm, g = np.zeros(2), np.zeros(6)
    st, Fh = np.zeros((27, N)), np.zeros((nsteps, 2))
    c = None if cot is None else np.asarray(cot, dtype=np.float64)
    assert L.adjh_chain(C.byref(cfg), nsteps, w0, w1, None if c is None else c.ctypes.data, Fm.ctypes.data,
                        g.ctypes.data if c is not None else None, st.ctypes.data, Fh.ctypes.data) == 0
    return Fm, (g if c is not None else None), st, Fh


def oracle_mean_force(cfg, nsteps, w0, w1):
    o = ob.OracleSim(cfg)
    o.step(nsteps)
    return o.force_history()[w0:w1].mean(axis=0)


@pytest.mark.parametrize("thermal", ["adiabatic", "contact", "pse"])
def test_host_chain_forward_is_the_oracle_step(thermal):
    cfg = adj_config(thermal)
    _, _, st, Fh = host_chain(cfg, 150, 50, 150)
    o = ob.OracleSim(cfg)
    o.step(150)
    Fo = o.force_history()
    assert np.abs(Fh - Fo).max() <= 1e-10 * np.abs(Fo).max()
    pos, S = o.field("pos"), o.field("S")
    assert np.abs(st[0] - pos[:, 0]).max() < 1e-14 and np.abs(st[1] - pos[:, 1]).max() < 1e-14
    assert np.abs(st[4] - S[:, 0]).max() <= 1e-9 * np.abs(S).max()
    assert np.abs(st[9] - o.field("T")).max() < 1e-8
    assert np.abs(st[10] - o.field("eps_pl")).max() < 1e-10
    assert o.field("eps_pl").max() > 1e-2  # the run is plastic


@pytest.mark.parametrize("thermal", ["adiabatic", "contact"])
def test_host_chain_gradient_matches_oracle_differences(thermal):
    cfg = adj_config(thermal)
    nsteps, w0, w1 = 150, 50, 150
    th = theta_of(cfg)
    for cot in ([1.0, 0.0], [0.0, 1.0]):
        _, g, _, _ = host_chain(cfg, nsteps, w0, w1, cot)

        def fd_at(rel):
            out = np.zeros(6)
            for j in range(6):
                h = rel * abs(th[j])
                tp, tm = th.copy(), th.copy()
                tp[j] += h
                tm[j] -= h
                out[j] = np.dot(cot, oracle_mean_force(with_theta(cfg, tp), nsteps, w0, w1) -
                                oracle_mean_force(with_theta(cfg, tm), nsteps, w0, w1)) / (2 * h)
            return out

        fd = fd_at(1e-6)
        unc = np.maximum(np.abs(fd - fd_at(1e-5)), np.abs(fd - fd_at(1e-7)))
        scale = np.abs(fd).max()
        err = np.abs(g - fd)
        print(thermal, cot, "grad", g, "\nfd  ", fd, "\nrel err", err / scale, "\nfd unc", unc / scale)
        assert (err <= 1e-4 * scale + 2 * unc).all()
        # not vacuous: entries where the differences agree to 1e-5 across step sizes meet 1e-4
        assert ((err < 1e-4 * scale) & (unc < 1e-5 * scale)).sum() >= 2


def test_host_chain_short_windows_from_a_plastic_state():
    """Steps 120..125 differentiated from the (fixed) state reached after 120 steps: central
    differences are clean there (no yield onset inside the window), so the bar is 1e-6."""
    cfg = adj_config("adiabatic")
    th = theta_of(cfg)
    pre, n = 120, 5

    def sub(t, cot=None):
        Fm, g = np.zeros(2), np.zeros(6)
        tt = np.asarray(t, dtype=np.float64)
        c = None if cot is None else np.asarray(cot, dtype=np.float64)
        L.adjh_chain_sub(C.byref(cfg), pre, pre + n, pre, pre + n, tt.ctypes.data,
                         None if c is None else c.ctypes.data, Fm.ctypes.data, g.ctypes.data if c is not None else None)
        return Fm, g

    _, g = sub(th, [1.0, 0.0])
    fd = np.zeros(6)
    for j in range(6):
        h = 1e-6 * abs(th[j])
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (sub(tp)[0][0] - sub(tm)[0][0]) / (2 * h)
    err = np.abs(g - fd) / np.abs(fd).max()
    assert (err < 1e-6).all(), (g, fd)
