"""The C ABI library loads and exports every function include/mfc.h declares; host-only
entry points work without a GPU."""
import ctypes as C
import os
import re

import numpy as np

import pymfc
import oracle_binding as ob

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "mfc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mfc_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_step_api():
    names = declared_functions()
    for n in ["mfc_config_ref_cut", "mfc_create", "mfc_step", "mfc_tool_force", "mfc_get_field", "mfc_get_cells",
              "mfc_destroy", "mfc_last_error", "mfc_profile", "mfc_scenario"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = pymfc.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_field_components():
    L = pymfc.lib()
    for name, (fid, nc) in pymfc.FIELDS.items():
        assert L.mfc_field_components(fid) == nc, name
    assert L.mfc_field_components(99) == 0


def test_config_ref_cut_literals():
    """setup_ref_cut's constants (benchmarks.cpp:266-332,384-421, mfree_gpu.cpp:160-165)."""
    c = pymfc.config_ref_cut()
    assert c.lattice == pymfc.LATTICE_REF and c.ny == 30
    assert (c.hi_x, c.hi_y, c.lo_x, c.hdx) == (0.2, 0.06, 0.005, 1.5)
    assert (c.E, c.nu, c.rho0) == (1.1, 0.35, 4.429998)
    assert (c.jc_A, c.jc_B, c.jc_C, c.jc_m, c.jc_n) == (0.00862, 0.00331, 0.01, 0.8, 0.34)
    assert (c.jc_Tref, c.jc_Tmelt, c.jc_eps_dot_ref) == (300.0, 1836.0, 1e-6)
    assert (c.cp, c.tq, c.k_wp, c.eta) == (553e-8, 0.9, 7.1e-13, 0.9)
    assert (c.rho_tool, c.cp_tool, c.k_tool) == (15.25, 292e-8, 88e-13)
    assert (c.av_alpha, c.av_beta, c.av_eta, c.xsph_eps, c.stress_eps) == (1.0, 1.0, 0.1, 0.5, 0.3)
    assert (c.rake_deg, c.clear_deg, c.edge_radius, c.feed) == (10.0, 8.0, 0.002, 0.01)
    assert c.speed == 83.333328e-5 / 500.0 * 70.0
    assert (c.mu, c.contact_alpha, c.sample_tool, c.cfl) == (0.35, 0.1, 1, 0.5)
    assert abs(c.tool_tl_x - (-0.046)) < 1e-15 and c.tool_tl_y == 0.0986074


def test_config_ref_cut_matches_oracle():
    for args in [(30, 10.0, 8.0, 0.002, 1.1666e-4, 0.01), (12, 0.0, 7.0, 0.0, 0.0, 0.02)]:
        a, b = pymfc.config_ref_cut(*args), ob.config_ref_cut(*args)
        assert bytes(a) == bytes(b)


def test_create_reports_errors_without_crashing():
    cfg = pymfc.config_ref_cut()
    cfg.thermal_solver = 7
    h = C.c_void_p()
    rc = pymfc.lib().mfc_create(C.byref(cfg), C.byref(h))
    assert rc != 0 and not h.value
    assert len(pymfc.lib().mfc_last_error()) > 0


def test_scenario_rejects_bad_lattice():
    cfg = pymfc.cut_config(10, 10, 10.0)
    cfg.dx = -1.0
    try:
        pymfc.scenario(cfg)
    except pymfc.MfcError as e:
        assert "lattice" in str(e)
    else:
        raise AssertionError("negative dx accepted")


def test_scenario_needs_no_gpu():
    s = pymfc.scenario(pymfc.cut_config(20, 10, 10.0))
    assert s["info"].n_particles == 200 and s["pos"].shape == (200, 2)
    assert np.all(np.isfinite(s["pos"]))


def test_load_refuses_a_file_that_is_not_a_checkpoint(tmp_path):
    """mfc_load checks the header before it touches a device."""
    p = tmp_path / "not_a_checkpoint.bin"
    p.write_bytes(b"MFC" + bytes(400))
    h = C.c_void_p()
    rc = pymfc.lib().mfc_load(os.fsencode(str(p)), C.byref(h))
    assert rc == -1 and not h.value
    assert b"not a checkpoint" in pymfc.lib().mfc_last_error()
    rc = pymfc.lib().mfc_load(os.fsencode(str(tmp_path / "missing.bin")), C.byref(h))
    assert rc == -1
