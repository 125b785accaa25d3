"""Parity of the HIP FP64 path with the FP64 host oracle, through the C ABI.

Bar: bit-exact (the FP64 kernels and the oracle evaluate the same expressions in the same
order with no FMA contraction and the same transcendental library); the required bound is
1e-10 relative in positions, stress, plastic strain, temperature and tool force."""
import numpy as np
import pytest

import pymfc
import oracle_binding as ob
from parity import assert_bitwise, compare, equal, max_rel

pytestmark = pytest.mark.gpu


def small_cut(**kw):
    """40 x 12 lattice at 10 um, tool on the edge: contact from step 1."""
    return pymfc.cut_config(40, 12, 10.0, **kw)


def run_pair(cfg, steps, chunks=1):
    g = pymfc.Sim(cfg)
    o = ob.OracleSim(cfg)
    per = steps // chunks
    for _ in range(chunks):
        g.step(per)
        o.step(per)
    return g, o


@pytest.mark.parametrize("fn,lo,hi,two", [
    (0, -745.0, 709.0, False), (1, 1e-300, 1e300, False), (2, 0.0, 2.0, True), (3, -5.0, 5.0, True),
    (4, -20.0, 20.0, False), (5, -20.0, 20.0, False), (6, -5.0, 5.0, True), (7, -2.0, 2.0, False),
    (8, -2.0, 2.0, False), (9, 0.0, 1e3, False), (10, 0.0, 1e3, False), (11, -1e3, 1e3, True),
    (12, -30.0, 30.0, False)])
def test_device_math_bitwise(fn, lo, hi, two):
    rng = np.random.default_rng(fn)
    if fn == 1:
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 200000))
    else:
        x = rng.uniform(lo, hi, 200000)
    y = rng.uniform(-3.0, 3.0, x.size) if two else None
    if fn == 2:
        y = rng.uniform(0.0, 3.0, x.size)
    specials = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-310, 0.5, 2.0])
    x = np.concatenate([x, specials])
    if y is not None:
        y = np.concatenate([y, specials[::-1]])
    d = pymfc.device_math(fn, x, y)
    h = ob.host_math(fn, x, y)
    assert equal(d, h), f"fn {fn}: {np.sum(~((d == h) | (np.isnan(d) & np.isnan(h))))} mismatches"


def test_binning_bitwise_first_step():
    cfg = pymfc.config_ref_cut(ny=12)
    g, o = run_pair(cfg, 1)
    assert equal(g.field("hash"), o.field("hash"))
    assert equal(g.field("id"), o.field("id"))
    assert equal(g.cells(), o.cells())


def test_binning_jittered_lattice():
    cfg = small_cut(jitter=0.3, seed=7)
    g, o = run_pair(cfg, 3)
    assert equal(g.field("hash"), o.field("hash"))
    assert equal(g.field("id"), o.field("id"))
    assert equal(g.cells(), o.cells())


# particles that have yielded (plastic strain > 0) at the end of each parity case, at least: the
# faithful plain Johnson-Cook return (float truncations, plastic heating, thermal softening) is
# exercised by the cuts at 1000 m/min
MIN_YIELDED = {"plastic_contact_1000": 60, "plastic_adiabatic_1000": 50, "tanh_jc": 40}


@pytest.mark.parametrize("case", ["ref_cut_thermal_tool", "adiabatic_edge", "pse_jitter", "sharp_tool",
                                  "contact_thermal", "promoted_tight_secant", "tanh_jc",
                                  "plastic_contact_1000", "plastic_adiabatic_1000"])
def test_step_parity(case):
    if case == "plastic_contact_1000":  # faithful plain JC in a thermal-contact cut that yields
        cfg = small_cut(thermal="contact", speed_m_min=1000.0)
        steps = 400
    elif case == "plastic_adiabatic_1000":
        cfg = small_cut(thermal="adiabatic", speed_m_min=1000.0)
        steps = 300
    elif case == "tanh_jc":  # the secant solves with the tanh-modified flow stress (Calamaz 2008)
        cfg = small_cut(thermal="contact", speed_m_min=1000.0)
        cfg.jc_model = pymfc.JC_TANH
        cfg.jc_tanh_a, cfg.jc_tanh_b, cfg.jc_tanh_c, cfg.jc_tanh_d = 2.0, 5.0, 1.0, 0.5
        steps = 200
    elif case == "promoted_tight_secant":
        cfg = small_cut(thermal="contact")
        cfg.promote_fp64 = 1
        cfg.secant_tol = 1e-13
        steps = 200
    elif case == "ref_cut_thermal_tool":
        cfg = pymfc.config_ref_cut(ny=12)
        steps = 40
    elif case == "adiabatic_edge":
        cfg = small_cut(thermal="adiabatic")
        steps = 300
    elif case == "pse_jitter":
        cfg = small_cut(thermal="pse", jitter=0.05, seed=11)
        steps = 150
    elif case == "sharp_tool":
        cfg = small_cut(thermal="adiabatic")
        cfg.edge_radius = 0.0
        cfg.rake_deg = 6.0
        steps = 200
    else:
        cfg = small_cut(thermal="contact")
        steps = 200
    g, o = run_pair(cfg, steps, chunks=2)
    assert_bitwise(g, o, context=case)
    assert equal(g.force_history(), o.force_history())
    assert equal(g.tool_force(), o.tool_force())
    assert g.info.out_of_grid == 0 and o.info.out_of_grid == 0
    yielded = int((o.field("eps_pl") > 0).sum())
    print(case, "yielded", yielded, "of", o.N)
    assert yielded >= MIN_YIELDED.get(case, 0)


def test_tanh_jc_changes_the_run():
    """The tanh model is on the path: a cut that yields differs from its plain-JC run, and the
    reverse mode, which differentiates plain JC, refuses it."""
    cfg = small_cut(thermal="contact", speed_m_min=1000.0)
    plain = pymfc.Sim(cfg)
    cfg.jc_model = pymfc.JC_TANH
    cfg.jc_tanh_a, cfg.jc_tanh_b, cfg.jc_tanh_c, cfg.jc_tanh_d = 2.0, 5.0, 1.0, 0.5
    tanh = pymfc.Sim(cfg)
    plain.step(100)
    tanh.step(100)
    assert plain.field("eps_pl").max() > 0
    assert not np.array_equal(plain.field("S"), tanh.field("S"))
    cfg.promote_fp64 = 1
    with pytest.raises(pymfc.MfcError):
        pymfc.Sim(cfg).vjp_forward(10, 5, 0, 10)


def test_graph_replay_matches_eager():
    cfg = small_cut(thermal="contact")
    g1 = pymfc.Sim(cfg)
    g1.step(64)  # graph path (captured graphs of 8 steps)
    g2 = pymfc.Sim(cfg)
    for _ in range(64):
        g2.step(1)  # eager path
    for f in ["pos", "vel", "S", "T", "eps_pl", "fc", "ft"]:
        assert equal(g1.field(f), g2.field(f)), f
    assert equal(g1.force_history(), g2.force_history())


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
@pytest.mark.parametrize("thermal", ["contact", "pse"])
@pytest.mark.parametrize("split,compact", [(1, 0), (0, 1), (1, 1)])
def test_interaction_variants_match(prec, thermal, split, compact):
    """Every launch variant of the interaction (split sums/update kernels, compacted pair
    loop) gives the bits of the fused kernel that walks the window directly."""
    cfg = small_cut(thermal=thermal, precision=prec)
    g1, g2 = pymfc.Sim(cfg), pymfc.Sim(cfg)
    g1.set_tuning(split_interact=0, compact_pairs=0)
    g2.set_tuning(split_interact=split, compact_pairs=compact)
    g1.step(30)
    g2.step(30)
    for f in ["pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags"]:
        assert equal(g1.field(f), g2.field(f)), f
    assert equal(g1.force_history(), g2.force_history())


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
@pytest.mark.parametrize("thermal", ["contact", "pse"])
@pytest.mark.parametrize("split,compact", [(0, 0), (1, 1)])
def test_unstaged_tiles_match(prec, thermal, split, compact):
    """A tile copies its window candidates into LDS when their sorted-index span fits the stage;
    otherwise it reads them from the records in memory. stage_limit=1 sends every tile down the
    second path: the results are the bits of the staged run (and so, in FP64, the oracle's)."""
    cfg = small_cut(thermal=thermal, precision=prec)
    g1, g2 = pymfc.Sim(cfg), pymfc.Sim(cfg)
    g1.set_tuning(split_interact=split, compact_pairs=compact)
    g2.set_tuning(split_interact=split, compact_pairs=compact, stage_limit=1)
    assert g2.tuning()["stage_limit"] == 1
    g1.step(30)
    g2.step(30)
    for f in ["pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags"]:
        assert equal(g1.field(f), g2.field(f)), f
    assert equal(g1.force_history(), g2.force_history())


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
@pytest.mark.parametrize("split", [0, 1])
def test_xcd_chunk_lengths_match(prec, split):
    """The interaction kernels deal their tiles to the XCDs in chunks of xcd_chunk tiles; the
    map decides only where a tile runs, so every chunk length gives the same bits (a cut of 55
    tiles: chunks of 1, 3, 32 and 64 tiles, the last one holding every tile)."""
    cfg = pymfc.cut_config(200, 60, 10.0, thermal="contact", speed_m_min=1000.0, precision=prec)
    runs = []
    for chunk in (0, 1, 3, 32, 64):
        g = pymfc.Sim(cfg)
        g.set_tuning(split_interact=split, xcd_chunk=chunk)
        assert g.tuning()["xcd_chunk"] == chunk
        g.step(30)
        runs.append(g)
    assert 48 * 256 < runs[0].N <= 64 * 256
    for g in runs[1:]:
        for f in ["pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags"]:
            assert equal(runs[0].field(f), g.field(f)), f
        assert equal(runs[0].force_history(), g.force_history())


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
def test_sort_paths_match(prec):
    """The counting sort takes a chunk's particles from its list into registers and LDS when the
    chunk holds at most sort_fast of them, walks the list with a bucket in memory when it holds
    more, and passes over every particle's cell id when it holds more than sort_list. Each cell's
    order is the rank by current position whatever the path, so every bound gives the same bits:
    sort_fast=1 sends every chunk of two or more particles down the list walk, sort_list=1 down
    the pass over all particles, and the bounds of 64 and 200 mix the three paths in one step."""
    cfg = pymfc.cut_config(200, 60, 10.0, thermal="contact", speed_m_min=1000.0, precision=prec)
    runs = []
    for fast, lst in ((0, 0), (1, 0), (0, 1), (64, 0), (64, 200)):
        g = pymfc.Sim(cfg)
        g.set_tuning(sort_fast=fast, sort_list=lst)
        t = g.tuning()
        assert (t["sort_fast"], t["sort_list"]) == (fast, lst)
        g.step(30)
        runs.append(g)
    for g in runs[1:]:
        for f in ["pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags", "id"]:
            assert equal(runs[0].field(f), g.field(f)), f
        assert equal(runs[0].cells(), g.cells())
        assert equal(runs[0].force_history(), g.force_history())


def test_sort_bounds_are_checked():
    g = pymfc.Sim(small_cut())
    for kw in ({"sort_fast": -1}, {"sort_fast": pymfc.SORT_FAST_MAX + 1}, {"sort_list": -1}):
        with pytest.raises(pymfc.MfcError):
            g.set_tuning(**kw)
    g.set_tuning(sort_fast=pymfc.SORT_FAST_MAX, sort_list=1)


def test_xcd_chunk_is_checked():
    g = pymfc.Sim(small_cut())
    for bad in (-1, pymfc.XCD_CHUNK_MAX + 1):
        with pytest.raises(pymfc.MfcError):
            g.set_tuning(xcd_chunk=bad)
    g.set_tuning(xcd_chunk=pymfc.XCD_CHUNK_MAX)


def test_stage_limit_is_checked():
    g = pymfc.Sim(small_cut())
    with pytest.raises(pymfc.MfcError):
        g.set_tuning(stage_limit=pymfc.STAGE_CAP + 1)
    with pytest.raises(pymfc.MfcError):
        g.set_tuning(stage_limit=-1)
    g.set_tuning(stage_limit=pymfc.STAGE_CAP)


def test_autotune_keeps_the_run():
    """The autotuner switches launch configurations while the run advances; every
    configuration computes the same bits, so the tuned run equals an untuned one."""
    cfg = small_cut(thermal="contact")
    a, b = pymfc.Sim(cfg), pymfc.Sim(cfg)
    t = pymfc.autotune(a, steps=5)
    assert len(t["ms_per_step"]) == 24
    assert {k: a.tuning()[k] for k in pymfc.TUNING_SPACE} == {k: t[k] for k in pymfc.TUNING_SPACE}
    assert t["steps_run"] == 2 * 24 * (16 + 5) + 16
    b.step(t["steps_run"])
    a.step(10)
    b.step(10)
    for f in ("pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags", "id"):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.force_history(), b.force_history())


def test_nan_blanking_matches_oracle():
    cfg = small_cut(thermal="contact")
    g, o = run_pair(cfg, 10)
    vt = o.field("vel_t")
    # a free workpiece particle: tool particles never read their velocity in the mechanics
    free = np.nonzero(o.field("flags") == 0)[0]
    vt[free[free.size // 2], 0] = np.nan
    g.set_field("vel_t", vt)
    o.set_field("vel_t", vt)
    g.step(20)
    o.step(20)
    fl = o.field("flags")
    assert (fl & pymfc.FLAG_BLANKED).sum() >= 1
    assert_bitwise(g, o, context="nan blanking")


def test_moved_particle_rebins():
    """A position written between steps is binned by the next step: the cell ids of a step's
    start state are computed by the previous step's interaction, so writing the state has to
    bin it again. The highest free particle is lifted by one cell, out of its cell."""
    cfg = small_cut(thermal="contact")
    g, o = run_pair(cfg, 10)
    pos, fl, ids = o.field("pos"), o.field("flags"), o.field("id")
    free = np.nonzero(fl == 0)[0]
    top = free[np.argmax(pos[free, 1])]
    before = o.field("hash")[top]
    pos[top, 1] += g.info.cell_size
    g.set_field("pos", pos)
    o.set_field("pos", pos)
    g.step(1)
    o.step(1)
    after = o.field("hash")[np.nonzero(o.field("id") == ids[top])[0][0]]
    assert after != before
    assert_bitwise(g, o, context="moved particle, first step")
    g.step(19)
    o.step(19)
    assert_bitwise(g, o, context="moved particle")


def test_adiabatic_6k_full_run():
    """The 150 x 40 FP64 adiabatic cut: 2000 steps, GPU vs oracle."""
    cfg = pymfc.workload("adiabatic_6k")
    g, o = run_pair(cfg, 2000, chunks=4)
    res = compare(g, o)
    bad = {f: r for f, r in res.items() if not r[0]}
    assert not bad, bad
    assert equal(g.force_history(), o.force_history())


# ----------------------------------------------------------------------------- FP32C accuracy
def fp32c_errors(cfg, steps):
    """Run the FP32-compensated path and the FP64 oracle side by side; return error measures."""
    c32 = cfg.copy()
    c32.precision = pymfc.PREC_FP32C
    g = pymfc.Sim(c32)
    o = ob.OracleSim(cfg)
    g.step(steps)
    o.step(steps)
    # compare per particle id (the FP32C run may sort ties differently once positions differ)
    gi, oi = np.argsort(g.field("id")), np.argsort(o.field("id"))
    dx = cfg.dx if cfg.dx > 0 else o.info.cell_size / cfg.hdx
    out = {}
    out["pos"] = np.abs(g.field("pos")[gi] - o.field("pos")[oi]).max() / dx
    S_o = o.field("S")[oi]
    out["S"] = np.abs(g.field("S")[gi] - S_o).max() / max(np.abs(S_o).max(), 1e-30)
    out["T"] = np.abs(g.field("T")[gi] - o.field("T")[oi]).max()
    out["Tpeak"] = abs(g.field("T").max() - o.field("T").max())
    out["eps_pl"] = np.abs(g.field("eps_pl")[gi] - o.field("eps_pl")[oi]).max()
    Fg, Fo = g.force_history(), o.force_history()
    w = max(1, steps // 2)
    mg, mo = Fg[-w:].mean(axis=0), Fo[-w:].mean(axis=0)
    out["F_mean_rel"] = np.abs(mg - mo) / np.maximum(np.abs(mo), 1e-30)
    out["blanked"] = int(np.count_nonzero(g.field("flags") & pymfc.FLAG_BLANKED))
    return out


@pytest.mark.parametrize("thermal", ["adiabatic", "contact", "pse"])
def test_fp32c_tracks_the_fp64_oracle(thermal):
    cfg = small_cut(thermal=thermal)
    e = fp32c_errors(cfg, 400)
    print("fp32c errors", thermal, e)
    assert e["blanked"] == 0
    assert e["pos"] < 1e-3          # positions within 1e-3 dx
    assert e["S"] < 2e-3            # stress within 0.2 % of the largest stress
    assert e["Tpeak"] < 1.0 and e["T"] < 1.0
    assert np.all(e["F_mean_rel"][0] < 5e-3)   # cutting force (time mean over the second half) within 0.5 %
