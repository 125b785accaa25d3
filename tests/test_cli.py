"""The command-line driver (mfree_iwf-ul_cut_gpu_amd/cli/mfc_cut.cpp): C++ host code over the C
ABI with the reference main()'s options, force log and wear log (mfree_gpu.cpp:146-323)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import pymfc

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "mfree_iwf-ul_cut_gpu_amd", "bin", "mfc_cut")


def _cli():
    if not os.path.exists(CLI):
        pytest.skip("mfc_cut not built")
    return CLI


def test_cli_usage():
    r = subprocess.run([_cli(), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "-r rake" in r.stderr


@pytest.mark.gpu
def test_cli_force_log_matches_the_library(tmp_path):
    # ny = 6: 249 particles, t_final/dt = 102106 steps, a force line every 756 steps
    r = subprocess.run([_cli(), "-n", "6", "-s", "1600", "-o", str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    logs = [f for f in os.listdir(tmp_path) if f.startswith("force_")]
    assert len(logs) == 1
    got = np.loadtxt(os.path.join(tmp_path, logs[0]))
    s = pymfc.Sim(pymfc.config_ref_cut(ny=6))
    s.step(1600)
    F = s.force_history()
    want = np.vstack([[0.0, 0.0], F[755], F[1511]])  # forces after 0, 756 and 1512 steps
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=0)
    # the wear log (mfree_gpu.cpp:238-299): min, max, mean and accumulated Usui wear at the same
    # steps, with the reference's constants and tool velocity (0, v) (benchmarks.cpp:508-512)
    wl = [f for f in os.listdir(tmp_path) if f.startswith("wear_")]
    assert len(wl) == 1
    W = np.loadtxt(os.path.join(tmp_path, wl[0]))
    assert W.shape == (3, 4)
    s2 = pymfc.Sim(pymfc.config_ref_cut(ny=6))
    dt = s2.info.dt
    want_w = []
    for n in (0, 756, 756):
        s2.step(n)
        want_w.append(s2.tool_wear(dt=756 * dt, K=1.0 / 100 * 7.8, alpha=2500.0))
    np.testing.assert_allclose(W, np.array(want_w), rtol=1e-5, atol=0)  # printed with %e
    # the VTK files of each dump (mfree_gpu.cpp:285-287)
    for q in range(3):
        for kind in ("out", "tool"):
            head = open(os.path.join(tmp_path, f"vtk_{kind}_{q:06d}.vtk")).readline()
            assert head.startswith("# vtk DataFile Version 2.0")


def test_cli_rejects_an_unknown_setup():
    r = subprocess.run([_cli(), "-b", "spheres"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "rings or impact" in r.stderr


@pytest.mark.gpu
def test_cli_runs_the_rings_setup(tmp_path):
    # -b rings -n 20: main()'s setup_rings line with nbox 20; no tool, so no tool VTK files
    r = subprocess.run([_cli(), "-b", "rings", "-n", "20", "-s", "40", "-o", str(tmp_path)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    s = pymfc.Sim(pymfc.config_rings(20))
    assert f"{s.N} particles (0 tool)" in r.stdout
    out = sorted(os.listdir(tmp_path))
    assert any(f.startswith("vtk_out_") for f in out) and not any(f.startswith("vtk_tool_") for f in out)
    F = np.loadtxt(os.path.join(tmp_path, [f for f in out if f.startswith("force_")][0]), ndmin=2)
    assert not F.any()
    s.step(40)
    head = open(os.path.join(tmp_path, "vtk_out_000000.vtk")).read().split("\n")
    assert head[4] == "DATASET UNSTRUCTURED_GRID" and head[5] == f"POINTS {s.N} float"
    s.close()
