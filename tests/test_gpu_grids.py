"""The reference's two cell grids on the device: grid_gpu_rothlin (MFC_GRID_ROTHLIN: cell ranges
rebuilt from the sorted ids by a separate pass, friction force not carried through the sort) bit
for bit against the oracle's restatement, and equal to the green grid where no friction history
exists."""
import numpy as np
import pytest

import pymfc
import oracle_binding as ob
from parity import assert_bitwise, equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64])
@pytest.mark.parametrize("split,compact", [(0, 1), (1, 0)])
def test_rothlin_cut_bitwise_with_oracle(prec, split, compact):
    cfg = pymfc.cut_config(40, 12, 10.0, thermal="contact", speed_m_min=1000.0, precision=prec)
    cfg.grid_kind = pymfc.GRID_ROTHLIN
    g, o = pymfc.Sim(cfg), ob.OracleSim(cfg)
    g.set_tuning(split_interact=split, compact_pairs=compact)
    for _ in range(2):
        g.step(100)
        o.step(100)
        assert equal(g.cells(), o.cells())
    assert_bitwise(g, o, context="rothlin cut")
    assert equal(g.force_history(), o.force_history())


def test_rothlin_equals_green_without_friction():
    """The rings have no tool, so no friction history: both grids give the same run, and the
    ranges of the rothlin pass equal those of the counting sort at every step checked."""
    runs = []
    for kind in (pymfc.GRID_GREEN, pymfc.GRID_ROTHLIN):
        cfg = pymfc.config_rings(30)
        cfg.grid_kind = kind
        runs.append(pymfc.Sim(cfg))
    for _ in range(5):
        for s in runs:
            s.step(13)
        assert equal(runs[0].cells(), runs[1].cells())
    for f in ("pos", "vel", "S", "rho", "hash", "id"):
        assert equal(runs[0].field(f), runs[1].field(f)), f


def test_rothlin_in_a_batch_equals_single():
    cfgs = [pymfc.cut_config(40, 12, 10.0, thermal="contact"), pymfc.cut_config(30, 10, 10.0, thermal="contact")]
    cfgs[1].grid_kind = pymfc.GRID_ROTHLIN
    b = pymfc.Batch(cfgs)
    b.step(20)
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        s.step(20)
        for f in ("pos", "vel", "S", "ft", "fc"):
            assert equal(b[i].field(f), s.field(f)), (i, f)
        s.close()
    b.close()
