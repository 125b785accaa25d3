"""The command-line driver (mfree_iwf-ul_cut_gpu_amd/cli/mfc_cut.cpp): C++ host code over the C
ABI with the reference main()'s options and force log (mfree_gpu.cpp:146-323)."""
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
