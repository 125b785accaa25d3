"""The committed fixtures of tests/golden/ (written by tests/golden/make_golden.py from the
oracle) are reproduced bit for bit: by the oracle on the CPU, by the HIP path on the GPU."""
import os

import numpy as np
import pytest

import pymfc
import oracle_binding as ob

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
import sys
sys.path.insert(0, GOLD)
import make_golden as mg  # noqa: E402


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_binning_fixture():
    g = load("binning_12x8_jitter.npz")
    o = ob.OracleSim(mg.cfg_binning())
    assert np.array_equal(o.field("pos"), g["pos0"])
    for s in (1, 2):
        o.step(1)
        assert np.array_equal(o.field("hash"), g[f"hash{s}"])
        assert np.array_equal(o.field("id"), g[f"id{s}"])
        assert np.array_equal(o.cells(), g[f"cells{s}"])


def _check_trace(sim):
    g = load("trace_20x10_tool.npz")
    assert sim.N == int(g["n_particles"])
    for s in range(1, 6):
        sim.step(1)
        for f in mg.TRACE_FIELDS:
            assert np.array_equal(sim.field(f), g[f"{f}_{s}"]), (s, f)
    assert np.array_equal(sim.force_history(), g["force"])


def test_trace_fixture_oracle():
    _check_trace(ob.OracleSim(mg.cfg_trace()))


@pytest.mark.gpu
def test_trace_fixture_gpu():
    _check_trace(pymfc.Sim(mg.cfg_trace()))


def test_jc_return_fixture():
    g = load("jc_return_table.npz")
    for a, b in zip(g["inputs"], g["outputs"]):
        out = ob.radial_return(g["jc"], float(g["G"]), float(g["dt"]), float(g["tq"]), float(g["cp"]), a)
        assert np.array_equal(out, b)


def test_contact_fixture():
    g = load("contact_cases.npz")
    o = ob.OracleSim(pymfc.workload("adiabatic_6k"))
    for p, v, f, want in zip(g["points"], g["vel"], g["ft_old"], g["out"]):
        assert np.array_equal(o.contact_point(p[0], p[1], v[0], v[1], f[0], f[1]), want)
