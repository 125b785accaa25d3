"""The reference's verification scenarios, setup_rings (benchmarks.cpp:155-253) and setup_impact
(benchmarks.cpp:23-153), in the oracle: the lattices and constants against closed forms, and
the runs against what the mechanics fixes — momentum of the free rings conserved to rounding,
the rings rebound (Gray, Monaghan and Swift 2001), the fixed wall rows held at zero velocity, plastic flow
in the impact and none when plasticity is switched off."""
import math

import numpy as np
import pytest

import pymfc
import oracle_binding as ob


def ring_lattice(nbox):
    """setup_rings' points by brute force: interleaved left/right images of the annulus."""
    ro, ri = 0.04, 0.03
    dx = 2 * ro / (nbox - 1)
    pts = []
    for i in range(nbox):
        for j in range(nbox):
            px, py = -ro + i * dx, -ro + j * dx
            d = math.sqrt(px * px + py * py)
            if ri <= d < ro:
                pts += [(px - (ro + 0.001), py), (px + (ro + 0.001), py)]
    return np.array(pts), dx


def test_rings_lattice_and_constants():
    cfg = pymfc.config_rings(30)
    o = ob.OracleSim(cfg)
    pts, dx = ring_lattice(30)
    assert np.array_equal(o.field("pos"), pts)
    info = o.info
    assert info.n_tool == 0 and info.n_particles == len(pts)
    assert abs(info.mass - dx * dx * 1.0) < 1e-20
    G = 1e7 / (2 * 1.4)
    K = 2 * G * 1.4 / (3 * (1 - 0.8))
    assert abs(info.dt - 0.5 * 1.7 * dx / (math.sqrt(K) + math.sqrt(170.0))) < 1e-18
    assert abs(info.t_final - 2.5 * 0.04 / 170.0) < 1e-18
    h = 1.7 * dx
    assert info.grid_nx == math.ceil((0.2 + 1e-6 + 0.2 + 1e-6) / h)
    assert info.grid_ny == math.ceil((0.08 + 1e-6 + 0.08 + 1e-6) / h)
    v = o.field("vel")
    assert np.array_equal(v[:, 0], np.where(pts[:, 0] < 0, 200.0, -200.0)) and not v[:, 1].any()
    assert not o.field("flags").any()
    assert len(o.tool()["left"]) == 0 and o.tool()["has_fillet"] == 0


def test_rings_conserve_momentum_and_rebound():
    cfg = pymfc.config_rings(30)
    o = ob.OracleSim(cfg)
    m = o.info.mass
    n = int(o.info.t_final / o.info.dt)
    scale = m * 200.0 * o.N
    left0 = o.field("id")[o.field("pos")[:, 0] < 0]
    for _ in range(4):
        o.step(n // 4)
        v = o.field("vel")
        assert abs(m * v[:, 0].sum()) < 1e-12 * scale
        assert abs(m * v[:, 1].sum()) < 1e-12 * scale
    ids, v = o.field("id"), o.field("vel")
    left = np.isin(ids, left0)
    assert v[left, 0].mean() < -50.0 and v[~left, 0].mean() > 50.0  # bounced back
    assert not o.field("eps_pl").any()     # elastic: no radial return
    assert not o.field("fc").any() and not o.field("ft").any()  # no tool, no contact
    assert o.out_of_grid() == 0
    assert np.isfinite(o.field("S")).all()


def test_impact_lattice_and_constants():
    nb = 20
    o = ob.OracleSim(pymfc.config_impact(nb))
    ro, dx = 0.2, 0.4 / (nb - 1)
    flyer = [(-ro + i * dx - (ro + 0.05), -ro + j * dx) for i in range(nb) for j in range(nb)
             if math.sqrt((-ro + i * dx) ** 2 + (-ro + j * dx) ** 2) < ro]
    wall = [(i * dx, j * dx - 1.0) for i in range(nb // 2) for j in range(10 * nb) if not j * dx - 1.0 > 1.0]
    pts = np.array(flyer + wall)
    assert np.array_equal(o.field("pos"), pts)
    v = o.field("vel")
    assert np.array_equal(v[:, 0], np.r_[np.full(len(flyer), 1e3), np.zeros(len(wall))])
    fixed = (o.field("flags") & pymfc.FLAG_FIXED) != 0
    assert np.array_equal(fixed, (pts[:, 1] < -1 + dx / 2) | (pts[:, 1] > 1 - dx / 2))
    assert fixed.sum() == 2 * (nb // 2)
    info = o.info
    G = 200e9 / 2.58
    K = 2 * G * 1.29 / (3 * (1 - 0.58))
    assert abs(info.dt - 0.3 * 1.7 * dx / (math.sqrt(K / 7830.0) + math.sqrt(1e3))) < 1e-18
    assert (o.field("T") == 273.0).all()


@pytest.mark.parametrize("elastic", [0, 1])
def test_impact_plastic_flow_and_fixed_rows(elastic):
    cfg = pymfc.config_impact(20)
    cfg.elastic = elastic
    o = ob.OracleSim(cfg)
    fixed0 = o.field("id")[(o.field("flags") & pymfc.FLAG_FIXED) != 0]
    n = int(o.info.t_final / o.info.dt)
    o.step(n // 2)
    ids, v = o.field("id"), o.field("vel")
    fx = np.isin(ids, fixed0)
    assert not v[fx].any()  # velocities held at zero (their positions still take the XSPH term)
    assert v[:, 0].max() < 1e3  # the flyer is slowed by the wall
    eps = o.field("eps_pl")
    if elastic:
        assert not eps.any()
    else:
        assert eps.max() > 0.1 and (eps > 0).sum() > 50
    assert o.out_of_grid() == 0 and np.isfinite(o.field("S")).all()
