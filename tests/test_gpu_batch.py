"""Batches of independent cutting instances: every instance of a batched run must equal its own
single-instance run bit for bit (forward, force history, reverse mode), for uneven sizes,
both precisions and both interaction variants."""
from __future__ import annotations

import numpy as np
import pytest

import pymfc
from parity import equal

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "S", "T", "eps_pl", "eps_pl_dot", "rho", "fc", "ft", "flags", "id", "hash")


def compare_instance(inst, single, what):
    for f in FIELDS:
        assert equal(inst.field(f), single.field(f)), f"{what}: field {f}"
    assert equal(inst.force_history(), single.force_history()), f"{what}: force history"
    assert equal(inst.tool_force(), single.tool_force()), f"{what}: tool force"


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
@pytest.mark.parametrize("split,compact", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_batch_equals_single_runs(prec, split, compact):
    cfgs = pymfc.batch_configs(8, seed=11, precision=prec)
    b = pymfc.Batch(cfgs)
    b.set_tuning(split_interact=split, compact_pairs=compact)
    b.step(17)  # two captured graphs of 8 steps plus one eager step
    b.step(3)
    b.sync()
    sizes = [inst.N for inst in b.instances]
    assert max(sizes) > 4 * min(sizes)  # uneven
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        s.step(20)
        compare_instance(b[i], s, f"instance {i} ({sizes[i]} particles)")
        s.close()
    b.close()


@pytest.mark.parametrize("prec", [pymfc.PREC_FP64, pymfc.PREC_FP32C])
def test_c5_batch_of_32_matches_single_runs(prec):
    cfgs = pymfc.batch_configs(32, precision=prec)
    b = pymfc.Batch(cfgs)
    b.step(20)
    b.sync()
    n = [inst.N for inst in b.instances]
    print("instances", len(n), "particles", sum(n), "min", min(n), "max", max(n))
    for i in range(len(cfgs)):  # every one of the 32 single runs
        s = pymfc.Sim(cfgs[i])
        s.step(20)
        compare_instance(b[i], s, f"instance {i}")
        assert b[i].info.out_of_grid == 0
        s.close()
    b.close()


def test_c5_batch_reverse_mode_of_32_equals_single_runs():
    """The reverse mode of all 32 instances of the seeded batch (promoted FP64, tight secant):
    mean forces and gradients equal each instance's own single run."""
    cfgs = []
    for c in pymfc.batch_configs(32):
        c.promote_fp64 = 1
        c.secant_tol = 1e-14
        cfgs.append(c)
    b = pymfc.Batch(cfgs)
    nsteps, K, w0, w1 = 40, 10, 10, 40
    F = b.vjp_forward(nsteps, K, w0, w1)
    rng = np.random.default_rng(32)
    cot = rng.normal(size=(32, 2))
    G = b.vjp_backward(cot)
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        Fi = s.vjp_forward(nsteps, K, w0, w1)
        Gi = s.vjp_backward(cot[i])
        assert equal(F[i], Fi), i
        assert equal(G[i], Gi), i
        s.close()
    b.close()


def test_batch_unstaged_equals_single_runs():
    """Every tile of every instance reading its candidates from memory (stage_limit=1) gives the
    bits of the single runs with the LDS stage."""
    cfgs = pymfc.batch_configs(4, seed=5, precision=pymfc.PREC_FP64)
    b = pymfc.Batch(cfgs)
    b.set_tuning(stage_limit=1)
    b.step(12)
    b.sync()
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        s.step(12)
        compare_instance(b[i], s, f"instance {i}")
        s.close()
    b.close()


def test_batch_reverse_mode_equals_single():
    cfgs = []
    for c in pymfc.batch_configs(3, seed=5):
        c.promote_fp64 = 1
        c.secant_tol = 1e-14
        cfgs.append(c)
    b = pymfc.Batch(cfgs)
    F = b.vjp_forward(30, 10, 10, 30)
    cot = np.array([[1.0, 0.5], [0.0, 1.0], [-1.0, 2.0]])
    G = b.vjp_backward(cot)
    for i, c in enumerate(cfgs):
        s = pymfc.Sim(c)
        Fi = s.vjp_forward(30, 10, 10, 30)
        Gi = s.vjp_backward(cot[i])
        assert equal(F[i], Fi), i
        assert equal(G[i], Gi), i
        s.close()
    b.close()


def test_batch_instances_refuse_single_steps():
    b = pymfc.Batch(pymfc.batch_configs(2, seed=3))
    with pytest.raises(pymfc.MfcError):
        b[0].step(1)
    b.close()
