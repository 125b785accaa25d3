"""Handle lifecycle on the GPU: repeated create / step (captured graphs) / per-phase profile /
change of launch configuration / destroy cycles in one process, in both precisions and both
interaction variants, each cycle's results equal to a reference run; and the out-of-grid count
of the binning."""
from __future__ import annotations

import numpy as np
import pytest

import pymfc
from parity import equal

pytestmark = pytest.mark.gpu


def test_create_step_profile_close_cycles():
    base = pymfc.cut_config(40, 12, 10.0, thermal="contact")
    ref = {}
    for prec in (pymfc.PREC_FP64, pymfc.PREC_FP32C):
        c = base.copy()
        c.precision = prec
        r = pymfc.Sim(c)
        r.step(20 + 3 + 9)
        ref[prec] = (r.field("pos"), r.field("T"), r.force_history())
        r.close()
    for cycle in range(8):
        prec = (pymfc.PREC_FP64, pymfc.PREC_FP32C)[cycle % 2]
        c = base.copy()
        c.precision = prec
        s = pymfc.Sim(c)
        s.set_tuning(split_interact=cycle // 2 % 2, compact_pairs=cycle // 4)
        s.step(20)                 # two graphs of 8 steps + 4 eager steps
        ph = s.profile(3)          # eager steps timed with events
        assert ph["interact"] > 0.0
        s.set_tuning(split_interact=1 - cycle // 2 % 2)  # drops the graphs
        s.step(9)                  # a new graph + one eager step
        pos, T, F = ref[prec]
        assert equal(s.field("pos"), pos), cycle
        assert equal(s.field("T"), T), cycle
        assert equal(s.force_history(), F), cycle
        s.close()


def test_out_of_grid_is_counted():
    """A particle moved outside the grid is counted by the next binning (its cell id is clamped
    into the table); a valid run counts none."""
    cfg = pymfc.cut_config(40, 12, 10.0, thermal="contact")
    s = pymfc.Sim(cfg)
    s.step(5)
    assert s.info.out_of_grid == 0
    pos = s.field("pos")
    fl = s.field("flags")
    k = int(np.flatnonzero(fl == 0)[0])
    info = s.info
    pos[k, 1] = info.bbmin_y + (info.grid_ny + 3) * info.cell_size  # above the grid
    s.set_field("pos", pos)
    s.step(1)
    assert s.info.out_of_grid >= 1
