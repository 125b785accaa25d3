"""Parity of the HIP path with the FP64 oracle at the benchmark sizes:

- the FP64 path bit for bit over a window of the thermal-contact cut at ~150k particles
  (750 x 200 at 2 um + 46,979 tool particles) and of the jittered PSE cut at 500k particles
  (1250 x 400 at 1.2 um);
- the FP32-compensated path against the oracle on the ~150k cut over a window the oracle runs in
  seconds: cutting and thrust force, peak temperature, stress, plastic strain and positions;
- the reverse mode at the calibration sizes (6,000 and 37,500 particles, a 400-step loss window,
  a checkpoint every 50 steps) against central differences of the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_binding as ob
import pymfc
from parity import assert_bitwise, equal
from test_gpu_adjoint import gpu_jacobian, oracle_fd_bracket, oracle_mean_force, theta_of, with_theta

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload,steps", [("thermal_contact_150k", 200), ("pse_500k", 100)])
def test_fp64_bitwise_at_full_size(workload, steps):
    cfg = pymfc.workload(workload, pymfc.PREC_FP64)
    g = pymfc.Sim(cfg)
    o = ob.OracleSim(cfg)
    g.step(steps)
    o.step(steps)
    print(workload, "N", g.N, "steps", steps, "yielded", int((o.field("eps_pl") > 0).sum()),
          "T max", o.field("T").max(), "F", o.tool_force())
    assert g.info.out_of_grid == 0 and o.info.out_of_grid == 0
    assert_bitwise(g, o, context=workload)
    assert equal(g.force_history(), o.force_history())


def test_fp32c_against_the_oracle_on_the_150k_cut():
    """Errors of the FP32-compensated path against the FP64 oracle after 500 steps of the
    thermal-contact cut (force: mean over the last 250 steps)."""
    steps = 500
    cfg = pymfc.workload("thermal_contact_150k", pymfc.PREC_FP64)
    o = ob.OracleSim(cfg)
    o.step(steps)
    g = pymfc.Sim(pymfc.workload("thermal_contact_150k", pymfc.PREC_FP32C))
    g.step(steps)
    gi, oi = np.argsort(g.field("id")), np.argsort(o.field("id"))
    Fg, Fo = g.force_history()[-steps // 2:].mean(axis=0), o.force_history()[-steps // 2:].mean(axis=0)
    S_o = o.field("S")[oi]
    e_o = o.field("eps_pl")[oi]
    err = {
        "force_rel": (np.abs(Fg - Fo) / np.abs(Fo)).tolist(),
        "Tpeak_K": float(abs(g.field("T").max() - o.field("T").max())),
        "T_max_abs_K": float(np.abs(g.field("T")[gi] - o.field("T")[oi]).max()),
        "S_max_rel": float(np.abs(g.field("S")[gi] - S_o).max() / np.abs(S_o).max()),
        "eps_pl_max_abs": float(np.abs(g.field("eps_pl")[gi] - e_o).max()),
        "eps_pl_max": float(e_o.max()),
        "pos_max_dx": float(np.abs(g.field("pos")[gi] - o.field("pos")[oi]).max() / 2e-4),
        "yielded": [int((g.field("eps_pl") > 0).sum()), int((e_o > 0).sum())],
        "F_oracle": Fo.tolist(),
    }
    print("fp32c vs oracle, thermal_contact_150k,", steps, "steps:", err)
    assert err["force_rel"][0] < 5e-3 and err["force_rel"][1] < 5e-3
    assert err["Tpeak_K"] < 1.0 and err["T_max_abs_K"] < 1.0
    assert err["S_max_rel"] < 2e-3
    assert err["pos_max_dx"] < 1e-3
    assert err["eps_pl_max_abs"] <= 1e-2 * max(err["eps_pl_max"], 1e-12)


def calibration_config(workload):
    cfg = pymfc.workload(workload, pymfc.PREC_FP64)
    cfg.promote_fp64 = 1
    cfg.secant_tol = 1e-14
    return cfg


@pytest.mark.parametrize("workload", ["adjoint_6k", "adjoint_40k"])
def test_calibration_gradient_at_full_size(workload):
    """The calibration loss |<F>(theta) - F*|^2 over a 400-step window with a checkpoint every 50
    steps; targets from JC constants perturbed by a seeded +-5 % and mu = 0.30.

    The Jacobian d<F>/dtheta of the reverse mode is checked against central differences of the
    oracle at a step of 1e-6 relative, bracketed by the steps 1e-5 and 1e-7. Over these 400 steps
    the tool advances about half a particle spacing into the workpiece, and the mean force hinges on
    the few particles of the contact zone switching between elastic and plastic. On the 6k cut
    (10 um spacing) the three difference estimates disagree by 10-5000 % on every entry: they do not
    resolve a derivative there, and the reverse mode is checked against the host restatement of
    the reverse chain instead (built from the same blocks, pairs pushed instead of pulled and the
    sums taken in another order, itself matched to central differences on cuts where those
    converge), to 1e-7 of its largest entry. On the 40k cut (4 um) the estimates agree to 1e-3 on
    the entries of n, C, m and mu, and there the reverse mode matches them to 1e-4; the A and B
    entries must lie within twice the estimates' spread."""
    cfg = calibration_config(workload)
    nsteps, K, w0, w1 = 400, 50, 0, 400
    F, J, sim = gpu_jacobian(cfg, nsteps, K, w0, w1)
    Fo = oracle_mean_force(cfg, nsteps, w0, w1)
    assert np.array_equal(F, Fo)
    fd, unc = oracle_fd_bracket(cfg, nsteps, w0, w1)
    err = np.abs(J - fd)
    spread = unc / np.maximum(np.abs(fd), 1e-300)
    print(workload, "N", sim.N, "F", F, "\nvjp", J, "\nfd ", fd, "\nfd spread (relative)", spread,
          "\nvjp - fd (relative)", err / np.maximum(np.abs(fd), 1e-300))
    if workload == "adjoint_40k":
        determinate = spread < 1e-3
        assert determinate.sum() >= 6
        assert (err[determinate] <= 1e-4 * np.abs(fd[determinate])).all()
        assert (err <= 1e-4 * np.abs(fd) + 2.0 * unc).all()
    else:
        from test_adjoint_chain import host_chain
        for c, cot in enumerate(([1.0, 0.0], [0.0, 1.0])):
            Fh, gh, _, _ = host_chain(cfg, nsteps, w0, w1, cot)
            assert np.abs(F - Fh).max() <= 1e-12 * np.abs(F).max()
            assert (np.abs(J[c] - gh) <= 1e-7 * np.abs(gh).max()).all(), (J[c], gh)
    # the loss gradient
    th0 = theta_of(cfg)
    rng = np.random.default_rng(0x355)
    th_t = th0 * (1.0 + rng.uniform(-0.05, 0.05, 6))
    th_t[5] = 0.30
    F_t = oracle_mean_force(with_theta(cfg, th_t), nsteps, w0, w1)
    g = sim.vjp_backward(2.0 * (F - F_t))
    print("loss", float(((F - F_t) ** 2).sum()), "grad", g, "from the differences", 2.0 * (F - F_t) @ fd)
    assert np.allclose(g, 2.0 * (F - F_t) @ J, rtol=1e-10, atol=1e-14 * np.abs(g).max())
