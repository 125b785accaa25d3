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


def test_impact_grid_follows_the_particles():
    """The impact's active grid (grid_gpu_green(10*n, n), benchmarks.cpp:135) is recomputed before
    every binning from the box of all positions and the largest h (grid.cu:97-136): after each
    step the geometry the oracle reports is the one of the positions at that step's start, the
    table holds 10 N cells, and every particle is inside it."""
    cfg = pymfc.config_impact(20)
    assert cfg.grid_adaptive == 1 and cfg.max_cell_factor == 10
    o = ob.OracleSim(cfg)
    info = o.info
    assert info.grid_adaptive == 1 and info.cell_table == 10 * o.N
    bbx = []
    for _ in range(6):
        p, h = o.field("pos"), o.field("h")
        o.step(10)  # the last of these steps bins the positions after nine more steps; recompute below
        bbx.append(o.info.bbmin_x)
    # one step from a known state: its geometry is that of the state's positions
    p, h = o.field("pos"), o.field("h")
    o.step(1)
    info = o.info
    x0, y0 = p[:, 0].min() - 1e-6, p[:, 1].min() - 1e-6
    cell = h.max() * 2.0 / 2.0
    nx = math.ceil(((p[:, 0].max() + 1e-6) - x0) / cell)
    ny = math.ceil(((p[:, 1].max() + 1e-6) - y0) / cell)
    assert (info.bbmin_x, info.bbmin_y, info.cell_size) == (x0, y0, cell)
    assert (info.grid_nx, info.grid_ny, info.num_cell) == (nx, ny, nx * ny)
    assert info.grid_overflow == 0 and o.out_of_grid() == 0
    assert bbx[-1] > bbx[0]  # the flyer moves right: the box's left edge follows it
    # the cell ranges cover every particle once, in the order of its cell
    off = o.cells()
    assert off[0] == 0 and off[-1] == o.N and np.all(np.diff(off) >= 0)
    hsh = o.field("hash")
    for c in np.unique(hsh)[:50]:
        assert np.all(hsh[off[c]:off[c + 1]] == c)


def test_impact_adaptive_and_locked_grids_agree():
    """The two grids sort the particles into different cells, so the neighbour sums add in a
    different order; the runs agree to rounding."""
    runs = []
    for adaptive in (True, False):
        o = ob.OracleSim(pymfc.config_impact(20, adaptive=adaptive))
        o.step(60)
        runs.append((o.field("id"), o.field("pos"), o.field("eps_pl")))
    (ia, pa, ea), (ib, pb, eb) = runs
    sa, sb = np.argsort(ia), np.argsort(ib)
    assert np.abs(pa[sa] - pb[sb]).max() < 1e-9
    assert np.abs(ea[sa] - eb[sb]).max() < 1e-9 and ea.max() > 0.05


def test_adaptive_grid_overflow_keeps_the_previous_geometry():
    """A particle thrown far out would need a grid larger than the 10 N cell table: the step
    keeps the previous geometry, counts the overflow, and counts the particle outside the grid.
    A first geometry that does not fit the table is refused at setup."""
    cfg = pymfc.config_impact(20)
    o = ob.OracleSim(cfg)
    o.step(3)
    before = o.info
    pos = o.field("pos")
    pos[0, 1] = 50.0
    o.set_field("pos", pos)
    o.step(1)
    after = o.info
    assert after.grid_overflow == 1 and after.out_of_grid == 1
    assert (after.bbmin_x, after.bbmin_y, after.grid_nx, after.grid_ny) == \
        (before.bbmin_x, before.bbmin_y, before.grid_nx, before.grid_ny)
    cfg.max_cell_factor = 1
    rings = pymfc.config_rings(30)
    rings.grid_adaptive, rings.max_cell_factor = 1, 1
    with pytest.raises(RuntimeError):
        ob.OracleSim(rings)
