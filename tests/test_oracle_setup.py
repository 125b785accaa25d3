"""Setup of the cut (setup_ref_cut, benchmarks.cpp:255-537; tool, tool.cpp:36-144,342-375):
geometry pinned by closed forms, and the solver library's host setup equal to the oracle's
bit for bit."""
import math

import numpy as np
import pytest

import pymfc
import oracle_binding as ob

CASES = {
    "ref_ny12": lambda: pymfc.config_ref_cut(ny=12),
    "ref_default": lambda: pymfc.config_ref_cut(),
    "ref_sharp": lambda: pymfc.config_ref_cut(ny=16, rake=5.0, clear=6.0, chamfer=0.0),
    "adiabatic_6k": lambda: pymfc.workload("adiabatic_6k"),
    "thermal_contact_150k": lambda: pymfc.workload("thermal_contact_150k"),
    "pse_500k": lambda: pymfc.workload("pse_500k"),
    "rings_80": lambda: pymfc.config_rings(80),
    "impact_20": lambda: pymfc.config_impact(20),
    "impact_57": lambda: pymfc.config_impact(57),
}


@pytest.mark.parametrize("name", list(CASES))
def test_library_setup_equals_oracle(name):
    cfg = CASES[name]()
    sc = pymfc.scenario(cfg)
    o = ob.OracleSim(cfg)
    for f in ("pos", "rho", "h", "T", "flags"):
        assert np.array_equal(sc[f], o.field(f)), f
    t1, t2 = sc["tool"], o.tool()
    for k in t1:
        assert np.array_equal(np.asarray(t1[k]), np.asarray(t2[k])), k
    assert sc["info"] == o.info


def test_lattice_and_constants_of_the_6k_cut():
    cfg = pymfc.workload("adiabatic_6k")
    sc = pymfc.scenario(cfg)
    info = sc["info"]
    assert info.n_particles == 6000 and info.n_tool == 0
    dx = 1e-3
    assert abs(info.mass - dx * dx * 4.43) < 1e-18
    K = 1.138 / (3 * (1 - 2 * 0.342))
    c0 = math.sqrt(K / 4.43)
    v = 70 / 60 * 1e-4
    assert abs(info.dt - 0.3 * 1.5 * dx / (c0 + math.sqrt(v))) < 1e-15
    p = sc["pos"]
    assert abs(p[:, 1].max() - 0.06) < 1e-15 and abs(p[:, 0].min() - 0.005) < 1e-15
    assert np.allclose(np.diff(np.unique(np.round(p[:, 0], 10))), dx)
    fixed = (sc["flags"] & pymfc.FLAG_FIXED) != 0
    bottom = p[:, 1] < p[:, 1].min() + 2.5 * dx
    assert np.all(fixed[bottom])              # three bottom rows
    left = p[:, 0] < p[:, 0].min() + 0.5 * dx
    assert np.all(fixed[left & (p[:, 1] < p[:, 1].min() + 0.5 * (0.06 - p[:, 1].min()))])
    assert np.all(fixed[p[:, 0] > p[:, 0].max() - dx]) and not np.any(fixed[(p[:, 1] > 0.05) & (p[:, 0] < 0.01)])


def test_tool_geometry():
    """Closed polygon with unit inward normals (inside: (left - q).n < 0 for every segment,
    tool.cpp:104-119), rake face at the rake angle from the vertical,
    clearance face at the clearance angle from the horizontal, fillet tangent to both faces,
    lowest point at the depth of cut, rake face on the workpiece edge."""
    cfg = pymfc.workload("adiabatic_6k")
    t = pymfc.scenario(cfg)["tool"]
    L, R, n = t["left"], t["right"], t["n"]
    assert len(L) == 5 and t["has_fillet"] == 1
    assert np.allclose(R, np.roll(L, -1, axis=0))
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0)
    cen = L.mean(axis=0)
    for k in range(5):
        mid = 0.5 * (L[k] + R[k])
        assert np.dot(mid - cen, n[k]) < 0
    rake_dir = R[1] - L[1]
    assert abs(math.degrees(math.atan2(abs(rake_dir[0]), abs(rake_dir[1]))) - 0.0) < 1e-9
    clear_dir = R[3] - L[3]
    assert abs(math.degrees(math.atan2(abs(clear_dir[1]), abs(clear_dir[0]))) - 7.0) < 1e-9
    cx, cy, r = t["fillet"][:3]
    assert abs(r - 5e-4) < 1e-18
    for k in (1, 3):  # distance of the fillet centre to the rake and clearance lines
        d = abs(np.cross(R[k] - L[k], np.array([cx, cy]) - L[k])) / np.linalg.norm(R[k] - L[k])
        assert abs(d - r) < 1e-12
    assert abs((cy - r) - (0.06 - 0.01)) < 1e-12
    assert abs(L[1][0] - 0.005) < 1e-12


def test_tool_particles_inside_tool():
    cfg = pymfc.config_ref_cut(ny=12)
    sc = pymfc.scenario(cfg)
    tool = (sc["flags"] & pymfc.FLAG_TOOL) != 0
    assert tool.sum() == sc["info"].n_tool > 0
    t = sc["tool"]
    P = sc["pos"][tool]
    inside_poly = np.ones(len(P), bool)
    for k in range(len(t["left"])):
        inside_poly &= ((t["left"][k] - P) @ t["n"][k]) < 0
    cx, cy, r = t["fillet"][:3]
    inside = inside_poly | (np.hypot(P[:, 0] - cx, P[:, 1] - cy) < r)
    assert inside.all()
    assert np.all(sc["rho"][tool] == 15.25)


def test_tool_update_is_incremental_translation():
    """tool.cpp:211-221: after k steps the end points moved by k dt v (incremental rounding)."""
    cfg = pymfc.config_ref_cut(ny=10)
    o = ob.OracleSim(cfg)
    t0 = o.tool()
    k = 25
    o.step(k)
    t1 = o.tool()
    dt = o.info.dt
    shift = k * dt * t0["vel"][0]
    assert np.allclose(t1["left"][:, 0] - t0["left"][:, 0], shift, rtol=1e-9, atol=0)
    assert np.array_equal(t1["left"][:, 1], t0["left"][:, 1])
    for q in range(len(t1["left"])):
        if not t1["vertical"][q]:
            y = t1["line_a"][q] * t1["left"][q][0] + t1["line_b"][q]
            assert abs(y - t1["left"][q][1]) < 1e-12
