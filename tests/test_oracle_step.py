"""The oracle's whole step against properties the mathematics fixes: pairwise momentum
conservation, SPH gradients of linear fields (closed-form stress rate and density rate),
Brookshaw and PSE Laplacians of a quadratic temperature, penalty contact and Coulomb
friction in closed form, NaN blanking, the tool-force reduction, and independence of the
OpenMP thread count."""
import numpy as np
import pytest

import pymfc
import oracle_binding as ob


def free_block(nx=30, ny=24, dx_um=10.0, thermal="adiabatic"):
    """A lattice with no fixed particles and the tool far away (no contact for many steps)."""
    c = pymfc.cut_config(nx, ny, dx_um, thermal=thermal, fixed_rows=0)
    c.fix_left = 0
    c.fix_right = 0
    c.sample_tool = 0
    c.tool_tl_x = -0.05  # front left of the workpiece: not moved onto the edge
    return c


def interior(o, margin_h=2.2):
    p = o.field("pos")
    h = o.field("h")[0]
    m = margin_h * h
    return ((p[:, 0] > p[:, 0].min() + m) & (p[:, 0] < p[:, 0].max() - m) &
            (p[:, 1] > p[:, 1].min() + m) & (p[:, 1] < p[:, 1].max() - m))


def test_momentum_is_conserved_by_the_pair_forces():
    o = ob.OracleSim(free_block())
    rng = np.random.default_rng(0)
    N = o.N
    S = rng.normal(0, 1e-3, (N, 4))
    v = rng.normal(0, 1e-5, (N, 2))
    rho = 4.43 * (1 + rng.normal(0, 1e-3, N))
    o.set_field("S", S)
    o.set_field("vel", v)
    o.set_field("rho", rho)
    o.step(1)
    vt = o.field("vel_t")
    tot = np.abs(vt).sum(axis=0)
    assert np.all(np.abs(vt.sum(axis=0)) < 1e-11 * tot)


@pytest.mark.parametrize("kind", ["shear", "stretch"])
def test_linear_velocity_field_gives_closed_form_rates(kind):
    o = ob.OracleSim(free_block())
    p = o.field("pos")
    a = 1e-3
    v = np.zeros_like(p)
    if kind == "shear":
        v[:, 0] = a * (p[:, 1] - p[:, 1].mean())
    else:
        v[:, 0] = a * (p[:, 0] - p[:, 0].mean())
    o.set_field("vel", v)
    inner = interior(o)
    ids = o.field("id")
    o.step(1)
    # map back: the step re-sorts; compare on the particles that were interior
    k = np.isin(o.field("id"), ids[inner])
    St = o.field("S_t")[k]
    rt = o.field("rho_t")[k]
    rho = o.field("rho")[k]
    G = 1.138 / (2 * 1.342)
    if kind == "shear":
        assert np.allclose(St[:, 1], G * a, rtol=0.03)
        assert np.all(np.abs(St[:, [0, 2, 3]]) < 1e-6 * G * a)
        assert np.all(np.abs(rt) < 1e-6 * a * rho)
    else:
        assert np.allclose(St[:, 0], 4 * G * a / 3, rtol=0.03)
        assert np.allclose(St[:, 2], -2 * G * a / 3, rtol=0.03)
        assert np.allclose(St[:, 3], -2 * G * a / 3, rtol=0.03)
        assert np.allclose(rt, -rho * a, rtol=0.03)


@pytest.mark.parametrize("thermal,tol", [("contact", 0.05), ("pse", 0.1)])
def test_quadratic_temperature_laplacian(thermal, tol):
    """T = T0 + c x^2: Brookshaw and PSE both give T_t = alpha lap T = 2 alpha c inside (PSE
    sums over the square 5 x 5 cell window, so its truncation depends on the particle's place
    in its cell: +-10 %)."""
    cfg = free_block(thermal=thermal)
    cfg.sample_tool = 0
    o = ob.OracleSim(cfg)
    p = o.field("pos")
    c = 1e4
    T = 300.0 + c * (p[:, 0] - p[:, 0].min()) ** 2
    o.set_field("T", T)
    inner = interior(o, 3.5)
    ids = o.field("id")
    o.step(1)
    k = np.isin(o.field("id"), ids[inner])
    alpha = cfg.k_wp / (cfg.rho0 * cfg.cp)
    Tt = o.field("T_t")[k]
    assert np.allclose(Tt, 2 * alpha * c, rtol=tol), (Tt.min(), Tt.max(), 2 * alpha * c)


def test_linear_temperature_has_no_brookshaw_flux_inside():
    cfg = free_block(thermal="contact")
    o = ob.OracleSim(cfg)
    p = o.field("pos")
    o.set_field("T", 300.0 + 5e3 * (p[:, 0] - p[:, 0].min()))
    inner = interior(o, 3.0)
    ids = o.field("id")
    o.step(1)
    k = np.isin(o.field("id"), ids[inner])
    alpha = cfg.k_wp / (cfg.rho0 * cfg.cp)
    h = 1.5e-3
    assert np.all(np.abs(o.field("T_t")[k]) < 1e-9 * alpha * 5e3 / h)


def test_contact_closed_form():
    cfg = pymfc.workload("adiabatic_6k")
    o = ob.OracleSim(cfg)
    info = o.info
    t = o.tool()
    face_x = t["left"][1][0]
    m, dt, alpha, mu = info.mass, info.dt, cfg.contact_alpha, cfg.mu
    d = 2e-5
    vt = t["vel"][0]
    # inside, rake face closest: fN = m d alpha / dt^2 along +x
    r = o.contact_point(face_x - d, 0.07, vt, 0.0, 0.0, 0.0)
    assert r[0] == 1.0
    assert abs(r[1] - m * d * alpha / dt ** 2) < 1e-9 * r[1] and r[2] == 0.0
    assert (r[5], r[6]) == (1.0, 0.0)
    # sliding along the face: stick while the trial force stays below mu |fN|
    vy = 1e-6
    r = o.contact_point(face_x - d, 0.07, vt, vy, 0.0, 0.0)
    fstar = -alpha * m * vy / dt
    assert abs(r[4] - fstar) < 1e-12 * abs(fstar) and r[3] == 0.0
    vy = 1.0
    r = o.contact_point(face_x - d, 0.07, vt, vy, 0.0, 0.0)
    assert abs(abs(r[4]) - mu * abs(r[1])) < 1e-12 * abs(r[4]) and r[4] < 0
    # history: the previous friction force enters the trial force
    r = o.contact_point(face_x - d, 0.07, vt, 0.0, 0.0, 1e-9)
    assert abs(r[4] - 1e-9) < 1e-24
    # outside: all zero
    r = o.contact_point(face_x + d, 0.07, 0.0, 0.0, 1.0, 1.0)
    assert np.all(r == 0.0)
    # under the fillet: normal radial from the fillet centre, depth r - |q - c|
    cx, cy, rr = t["fillet"][:3]
    q = np.array([cx + 0.3 * rr, cy - 0.8 * rr])
    r = o.contact_point(q[0], q[1], vt, 0.0, 0.0, 0.0)
    nrm = (q - [cx, cy]) / np.linalg.norm(q - [cx, cy])
    assert r[0] == 1.0 and np.allclose(r[5:7], nrm, rtol=1e-12)
    depth = rr - np.linalg.norm(q - [cx, cy])
    assert abs(np.hypot(r[1], r[2]) - m * depth * alpha / dt ** 2) < 1e-9 * np.hypot(r[1], r[2])


def test_blanked_particle_is_frozen_and_ignored():
    cfg = pymfc.cut_config(30, 12, 10.0, thermal="contact")
    o = ob.OracleSim(cfg)
    o.step(5)
    ids = o.field("id")
    fl = o.field("flags")
    w = np.flatnonzero((fl & pymfc.FLAG_TOOL) == 0)[100]  # a workpiece particle
    victim = ids[w]
    fl[w] |= pymfc.FLAG_BLANKED
    o.set_field("flags", fl)
    p0 = o.field("pos")[w].copy()
    o.step(10)
    k = np.flatnonzero(o.field("id") == victim)[0]
    assert o.field("flags")[k] & pymfc.FLAG_BLANKED
    assert np.array_equal(o.field("pos")[k], p0)
    assert np.count_nonzero(o.field("flags") & pymfc.FLAG_BLANKED) == 1


def test_nan_derivative_blanks_the_particle():
    cfg = pymfc.cut_config(30, 12, 10.0, thermal="contact")
    o = ob.OracleSim(cfg)
    o.step(5)
    w = np.flatnonzero((o.field("flags") & pymfc.FLAG_TOOL) == 0)[150]  # a workpiece particle
    victim = o.field("id")[w]
    vt = o.field("vel_t")
    vt[w, 1] = np.nan
    o.set_field("vel_t", vt)
    pos0 = dict(zip(o.field("id"), o.field("pos")))
    o.step(1)
    fl = o.field("flags")
    ids = o.field("id")
    k = np.flatnonzero(ids == victim)[0]
    assert fl[k] & pymfc.FLAG_BLANKED
    # the NaN reaches the neighbours of the step it was injected in (their pair terms), no further
    blanked = np.flatnonzero(fl & pymfc.FLAG_BLANKED)
    h = o.field("h")[0]
    d = [np.linalg.norm(pos0[ids[b]] - pos0[victim]) for b in blanked]
    assert np.all(np.array(d) < 4 * h)
    nb = len(blanked)
    o.step(5)
    assert np.count_nonzero(o.field("flags") & pymfc.FLAG_BLANKED) == nb


def test_tool_force_reduction():
    cfg = pymfc.workload("adiabatic_6k")
    o = ob.OracleSim(cfg)
    o.step(60)
    fc, ft = o.field("fc"), o.field("ft")
    F = o.tool_force()
    want = fc.sum(axis=0) + ft.sum(axis=0)
    assert np.all(np.abs(F - want) <= 1e-12 * (np.abs(fc).sum(axis=0) + np.abs(ft).sum(axis=0)))
    assert F[0] > 0 and np.array_equal(o.force_history()[-1], F)


def test_thread_count_does_not_change_results():
    cfg = pymfc.cut_config(40, 16, 10.0, thermal="contact")
    n0 = ob.threads()
    try:
        ob.set_threads(1)
        a = ob.OracleSim(cfg)
        a.step(30)
        ob.set_threads(max(n0, 4))
        b = ob.OracleSim(cfg)
        b.step(30)
    finally:
        ob.set_threads(n0)
    for f in ("pos", "vel", "S", "T", "eps_pl", "fc", "ft", "id"):
        assert np.array_equal(a.field(f), b.field(f)), f
