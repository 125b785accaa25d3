"""Binary checkpoints (mfc_save / mfc_load): a loaded handle continues bit for bit as the run it
was saved from, in both precisions and for an instance of a batch; damaged files are refused."""
from __future__ import annotations

import os

import numpy as np
import pytest

import pymfc
from parity import equal

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R", "flags", "id", "hash",
          "pos_t", "vel_t", "rho_t", "S_t", "T_t")


def same_run(a, b, what):
    for f in FIELDS:
        assert equal(a.field(f), b.field(f)), f"{what}: field {f}"
    assert equal(a.force_history(), b.force_history()), f"{what}: force history"
    assert equal(a.tool_force(), b.tool_force()), f"{what}: tool force"
    assert a.info.step == b.info.step


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
@pytest.mark.parametrize("thermal", ["contact", "pse"])
def test_save_load_continue(prec, thermal, tmp_path):
    cfg = pymfc.cut_config(40, 12, 10.0, thermal=thermal, speed_m_min=1000.0, precision=prec)
    a = pymfc.Sim(cfg)
    a.set_tuning(compact_pairs=1)
    a.step(37)
    path = str(tmp_path / "cut.mfc")
    a.save(path)
    a.step(45)
    b = pymfc.Sim.load(path)
    assert b.tuning() == a.tuning()
    assert b.cfg.as_dict() == cfg.as_dict()
    b.step(45)
    same_run(a, b, f"prec {prec} {thermal}")
    assert a.field("eps_pl").max() > 0  # the continued run is in the plastic regime


def test_batch_instance_checkpoint(tmp_path):
    cfgs = pymfc.batch_configs(4, seed=9)
    bt = pymfc.Batch(cfgs)
    bt.step(12)
    path = str(tmp_path / "inst2.mfc")
    bt[2].save(path)
    bt.step(10)
    s = pymfc.Sim.load(path)
    s.step(10)
    same_run(bt[2], s, "batch instance 2")
    bt.close()


def test_damaged_checkpoints_are_refused(tmp_path):
    a = pymfc.Sim(pymfc.cut_config(20, 8, 10.0, thermal="contact"))
    a.step(5)
    path = str(tmp_path / "c.mfc")
    a.save(path)
    raw = open(path, "rb").read()
    bad = str(tmp_path / "bad.mfc")
    with open(bad, "wb") as f:  # one payload byte flipped
        f.write(raw[:-100] + bytes([raw[-100] ^ 1]) + raw[-99:])
    with pytest.raises(pymfc.MfcError, match="checksum"):
        pymfc.Sim.load(bad)
    with open(bad, "wb") as f:
        f.write(raw[: len(raw) // 2])
    with pytest.raises(pymfc.MfcError, match="truncated"):
        pymfc.Sim.load(bad)
    os.remove(bad)
    with pytest.raises(pymfc.MfcError):
        pymfc.Sim.load(bad)
