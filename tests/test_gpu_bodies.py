"""The reference's verification scenarios on the device: setup_rings (benchmarks.cpp:155-253)
and setup_impact (benchmarks.cpp:23-153) run through the same step kernels as the cut, with no
tool and, for the rings, no plasticity. FP64 is bit-identical to the oracle; FP32C keeps the
rings' momentum and rebound; a batch runs them beside a cut."""
import numpy as np
import pytest

import pymfc
import oracle_binding as ob
from parity import assert_bitwise, equal

pytestmark = pytest.mark.gpu

CASES = {
    "rings_30": (lambda p: pymfc.config_rings(30, p), 400),
    "impact_20": (lambda p: pymfc.config_impact(20, p), 120),                   # adaptive grid (active line)
    "impact_20_locked": (lambda p: pymfc.config_impact(20, p, adaptive=False), 120),  # the commented fixed grid
}


def by_id(sim, f):
    return sim.field(f)[np.argsort(sim.field("id"), kind="stable")]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("split,compact", [(0, 1), (1, 0)])
def test_bodies_bitwise_with_oracle(name, split, compact):
    make, steps = CASES[name]
    cfg = make(pymfc.PREC_FP64)
    print(f"case {name} split={split} compact={compact}", flush=True)
    g, o = pymfc.Sim(cfg), ob.OracleSim(cfg)
    g.set_tuning(split_interact=split, compact_pairs=compact)
    assert equal(by_id(g, "vel"), by_id(o, "vel"))  # initial velocities uploaded
    for _ in range(2):
        g.step(steps // 2)
        o.step(steps // 2)
    assert_bitwise(g, o, context=name)
    assert equal(g.force_history(), o.force_history())
    assert not np.any(g.tool_force())
    gi, oi = g.info, o.info
    for f in ("grid_nx", "grid_ny", "num_cell", "bbmin_x", "bbmin_y", "cell_size", "cell_table", "grid_adaptive",
              "out_of_grid", "grid_overflow"):
        assert getattr(gi, f) == getattr(oi, f), f
    assert gi.out_of_grid == 0 and gi.grid_overflow == 0
    assert equal(g.cells(), o.cells())
    g.close()


def test_adaptive_grid_overflow_matches_oracle():
    """A particle thrown far out: the geometry pass keeps the previous grid (the cell table of
    10 N cells cannot hold the new one), counts the overflow and the particle outside the grid,
    on the device as in the oracle."""
    cfg = pymfc.config_impact(20)
    g, o = pymfc.Sim(cfg), ob.OracleSim(cfg)
    g.step(3)
    o.step(3)
    pos = o.field("pos")
    pos[0, 1] = 50.0
    g.set_field("pos", pos)
    o.set_field("pos", pos)
    g.step(5)
    o.step(5)
    assert_bitwise(g, o, context="overflow")
    gi, oi = g.info, o.info
    assert gi.grid_overflow == oi.grid_overflow == 5
    assert gi.out_of_grid == oi.out_of_grid == 5
    assert (gi.grid_nx, gi.grid_ny, gi.bbmin_x) == (oi.grid_nx, oi.grid_ny, oi.bbmin_x)


def test_rings_fp32c_keeps_momentum_and_rebound():
    runs = {}
    for p in (pymfc.PREC_FP64, pymfc.PREC_FP32C):
        s = pymfc.Sim(pymfc.config_rings(40, p))
        x0 = s.field("pos")[:, 0]
        left = s.field("id")[x0 < 0]  # the left ring (before the first step: creation order)
        n = int(s.info.t_final / s.info.dt)
        s.step(n)
        ids, v = s.field("id"), s.field("vel")
        m = s.info.mass
        runs[p] = dict(px=m * v[:, 0].sum(), scale=m * 200.0 * s.N, vl=v[np.isin(ids, left), 0].mean(),
                       ke=0.5 * m * (v ** 2).sum())
        s.close()
    a, b = runs[pymfc.PREC_FP64], runs[pymfc.PREC_FP32C]
    assert abs(a["px"]) < 1e-12 * a["scale"] and abs(b["px"]) < 1e-4 * b["scale"]
    assert a["vl"] < -50.0 and b["vl"] < -50.0
    assert abs(b["vl"] - a["vl"]) < 0.02 * abs(a["vl"])
    # the kinetic energy left after the collision (ring vibration included): 2.2 % apart at
    # 40 lattice points across, 1381 steps
    assert abs(b["ke"] - a["ke"]) < 0.05 * a["ke"]


def test_bodies_in_a_batch_beside_a_cut():
    """A batch mixing locked grids (the rings, the cut) and the impact's adaptive grid."""
    cfgs = [pymfc.config_rings(24), pymfc.config_impact(16), pymfc.cut_config(40, 12, 10.0, thermal="contact"),
            pymfc.config_impact(14, adaptive=False)]
    b = pymfc.Batch(cfgs)
    b.step(19)
    b.sync()
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        s.step(19)
        for f in ("pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft"):
            assert equal(b[i].field(f), s.field(f)), f"instance {i}: field {f}"
        assert equal(b[i].tool_force(), s.tool_force())
        s.close()
    b.close()


def test_reverse_mode_refuses_the_bodies():
    for make in (pymfc.config_rings, pymfc.config_impact):
        cfg = make(16)
        cfg.promote_fp64 = 1
        s = pymfc.Sim(cfg)
        with pytest.raises(pymfc.MfcError):
            s.vjp_forward(10, 5, 0, 10)
        s.close()
