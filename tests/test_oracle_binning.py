"""Cell assignment and the stable sort of the oracle against brute force (grid.cu:34-51,
grid_gpu_green.cu:177-201): the order after each step is the stable sort by cell of the previous
order, and the cell table of the green grid (start/end) equals the offsets table of the
rothlin grid (grid_gpu_rothlin.cu:28-127)."""
import numpy as np

import pymfc
import oracle_binding as ob


def brute_hash(pos, info):
    ix = np.trunc((pos[:, 0] - info.bbmin_x) / info.cell_size).astype(np.int64)
    iy = np.trunc((pos[:, 1] - info.bbmin_y) / info.cell_size).astype(np.int64)
    return ix * info.grid_ny + iy


def brute_offsets(sorted_hash, num_cell):
    return np.searchsorted(sorted_hash, np.arange(num_cell + 1), side="left").astype(np.int32)


def check_step(o, pos_before, id_before, blanked_before, hash_before):
    info = o.info
    h = brute_hash(pos_before, info)
    h = np.where(blanked_before, hash_before, h)  # blanked particles keep their last cell id
    order = np.argsort(h, kind="stable")
    o.step(1)
    assert np.array_equal(o.field("id"), id_before[order])
    assert np.array_equal(o.field("hash"), h[order])
    assert np.array_equal(o.cells(), brute_offsets(h[order], info.num_cell))


def test_stable_sort_on_jittered_lattice():
    cfg = pymfc.cut_config(12, 8, 10.0, jitter=0.45, seed=3)
    o = ob.OracleSim(cfg)
    for _ in range(3):
        fl = o.field("flags")
        check_step(o, o.field("pos"), o.field("id"), (fl & 2) != 0, o.field("hash"))


def test_ties_keep_previous_order():
    """Particles sharing a cell keep their relative order from the previous step."""
    cfg = pymfc.cut_config(30, 20, 10.0, jitter=0.2, seed=5)
    o = ob.OracleSim(cfg)
    o.step(1)
    ids1, h1 = o.field("id"), o.field("hash")
    o.step(1)
    ids2, h2 = o.field("id"), o.field("hash")
    rank1 = np.empty_like(ids1)
    rank1[ids1] = np.arange(ids1.size)
    for c in np.unique(h2):
        members = ids2[h2 == c]
        assert np.all(np.diff(rank1[members]) > 0)


def test_blanked_particles_keep_stale_cell():
    cfg = pymfc.cut_config(20, 10, 10.0)
    o = ob.OracleSim(cfg)
    o.step(1)
    fl = o.field("flags")
    fl[[3, 50, 77]] |= pymfc.FLAG_BLANKED
    o.set_field("flags", fl)
    pos = o.field("pos")
    pos[3] += 0.01  # moved far away: a live particle would change cell
    o.set_field("pos", pos)
    fl = o.field("flags")
    check_step(o, o.field("pos"), o.field("id"), (fl & 2) != 0, o.field("hash"))


def test_green_start_end_equals_rothlin_offsets():
    """offsets[c]..offsets[c+1] of every occupied cell are its [start, end); empty cells are empty."""
    cfg = pymfc.config_ref_cut(ny=12)
    o = ob.OracleSim(cfg)
    o.step(2)
    off = o.cells()
    h = o.field("hash")
    starts = np.flatnonzero(np.r_[True, h[1:] != h[:-1]])
    ends = np.r_[starts[1:], h.size]
    for s, e in zip(starts, ends):
        assert off[h[s]] == s and off[h[s] + 1] == e
    occupied = np.zeros(o.info.num_cell, bool)
    occupied[h] = True
    assert np.all((np.diff(off) > 0) == occupied)


def test_hash_truncates_toward_zero():
    """(int)((x - x0)/dx) truncates: cell column 0 also takes the sliver below bbmin."""
    cfg = pymfc.cut_config(10, 6, 10.0)
    o = ob.OracleSim(cfg)
    info = o.info
    pos = o.field("pos")
    pos[0] = [info.bbmin_x - 0.4 * info.cell_size, info.bbmin_y + 2.5 * info.cell_size]
    o.set_field("pos", pos)
    ids = o.field("id")
    o.step(1)
    k = np.flatnonzero(o.field("id") == ids[0])[0]
    assert o.field("hash")[k] == 2


def test_out_of_grid_counted_per_axis():
    """A particle above the grid gets the reference's id ix * ny + iy, which lands in the next
    column (inside the cell table) — it is still counted as outside the grid."""
    import pymfc
    cfg = pymfc.cut_config(40, 12, 10.0, thermal="contact")
    o = ob.OracleSim(cfg)
    o.step(2)
    assert o.info.out_of_grid == 0
    info = o.info
    pos = o.field("pos")
    k = int(np.flatnonzero(o.field("flags") == 0)[0])
    pos[k, 1] = info.bbmin_y + (info.grid_ny + 0.5) * info.cell_size
    ix = int((pos[k, 0] - info.bbmin_x) / info.cell_size)
    assert 0 <= ix * info.grid_ny + info.grid_ny < info.num_cell  # the id itself stays in the table
    o.set_field("pos", pos)
    o.step(1)
    assert o.info.out_of_grid == 1


def test_rothlin_ranges_and_friction_history():
    """grid_gpu_rothlin (grid_gpu_rothlin.cu:28-250): the unique / scatter / scan ranges equal the
    green grid's; without friction (the rings) the runs are identical, in a cut the friction force
    left in place by its sort changes the run."""
    import pymfc
    for make, differs in ((lambda: pymfc.config_rings(24), False),
                          (lambda: pymfc.cut_config(40, 12, 10.0, thermal="contact", speed_m_min=1000.0), True)):
        runs = []
        for kind in (pymfc.GRID_GREEN, pymfc.GRID_ROTHLIN):
            cfg = make()
            cfg.grid_kind = kind
            o = ob.OracleSim(cfg)
            o.step(60)
            runs.append(o)
        g, r = runs
        if not differs:
            for f in ("pos", "vel", "S", "ft", "hash", "id"):
                assert np.array_equal(g.field(f), r.field(f)), f
            assert np.array_equal(g.cells(), r.cells())
        else:
            assert not np.array_equal(g.field("ft"), r.field("ft"))
        # the rothlin ranges of its own state are the counting-sort ranges of that state
        h = r.field("hash")
        off = r.cells()
        occupied = np.zeros(r.info.num_cell, bool)
        occupied[h] = True
        assert np.all((np.diff(off) > 0) == occupied)
        starts = np.flatnonzero(np.r_[True, h[1:] != h[:-1]])
        assert np.array_equal(off[h[starts]], starts)
