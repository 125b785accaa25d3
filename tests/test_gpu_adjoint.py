"""Reverse mode on the GPU: d(mean tool force)/d(theta), theta = (JC A, B, n, C, m, friction mu),
against central differences of the FP64 oracle (step 1e-6 relative, the smooth promoted map
with a tight secant tolerance), plus the torch.autograd.Function wrapper used for calibration."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_binding as ob
import pymfc

pytestmark = pytest.mark.gpu


def adj_config(thermal="adiabatic", nx=40, ny=12, dx=10.0):
    # 1000 m/min: the contact zone yields within 50 steps, so every parameter matters
    cfg = pymfc.cut_config(nx, ny, dx, thermal=thermal, speed_m_min=1000.0)
    cfg.promote_fp64 = 1
    cfg.secant_tol = 1e-14
    return cfg


def theta_of(cfg) -> np.ndarray:
    return np.array([getattr(cfg, n) for n in pymfc.THETA])


def with_theta(cfg, th):
    c = cfg.copy()
    for n, v in zip(pymfc.THETA, th):
        setattr(c, n, float(v))
    return c


def oracle_mean_force(cfg, nsteps, w0, w1) -> np.ndarray:
    o = ob.OracleSim(cfg)
    o.step(nsteps)
    return o.force_history()[w0:w1].mean(axis=0)


def oracle_fd(cfg, nsteps, w0, w1, rel=1e-6) -> np.ndarray:
    """J[c, j] = d Fmean_c / d theta_j by central differences of the oracle"""
    th = theta_of(cfg)
    J = np.zeros((2, 6))
    for j in range(6):
        h = rel * abs(th[j])
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        J[:, j] = (oracle_mean_force(with_theta(cfg, tp), nsteps, w0, w1) -
                   oracle_mean_force(with_theta(cfg, tm), nsteps, w0, w1)) / (2 * h)
    return J


def oracle_fd_bracket(cfg, nsteps, w0, w1):
    """FD at steps 1e-6 (the reference value) and its uncertainty: the spread against steps 1e-5
    and 1e-7. Elastic/plastic and stick/slip switches of single particles make the mean force
    piecewise smooth; where a switch falls inside the difference stencil the three estimates
    disagree, and that disagreement bounds what central differences can resolve there."""
    f6 = oracle_fd(cfg, nsteps, w0, w1, 1e-6)
    f5 = oracle_fd(cfg, nsteps, w0, w1, 1e-5)
    f7 = oracle_fd(cfg, nsteps, w0, w1, 1e-7)
    return f6, np.maximum(np.abs(f6 - f5), np.abs(f6 - f7))


def gpu_jacobian(cfg, nsteps, K, w0, w1):
    sim = pymfc.Sim(cfg)
    F = sim.vjp_forward(nsteps, K, w0, w1)
    J = np.stack([sim.vjp_backward([1.0, 0.0]), sim.vjp_backward([0.0, 1.0])])
    return F, J, sim


def assert_jacobian(J, fd, unc, tol=1e-4):
    # per force component, relative to the largest sensitivity of that component (parameters
    # the force barely depends on are held to the same absolute accuracy), widened by twice
    # the differences' own uncertainty where a switch falls inside the stencil
    smooth = 0
    for c in range(2):
        scale = np.abs(fd[c]).max()
        err = np.abs(J[c] - fd[c])
        print(f"component {c}: rel err {err / scale}, fd uncertainty {unc[c] / scale}")
        assert (err <= tol * scale + 2 * unc[c]).all(), f"component {c}: rel err {err / scale}\nvjp {J[c]}\nfd  {fd[c]}"
        smooth += int(((unc[c] < 1e-5 * scale) & (err < tol * scale)).sum())
    assert smooth >= 4  # entries where the differences agree to 1e-5 across step sizes meet 1e-4


@pytest.mark.parametrize("thermal", ["adiabatic", "contact"])
def test_vjp_matches_oracle_differences(thermal):
    cfg = adj_config(thermal)
    nsteps, K, w0, w1 = 150, 25, 50, 150
    F, J, _ = gpu_jacobian(cfg, nsteps, K, w0, w1)
    Fo = oracle_mean_force(cfg, nsteps, w0, w1)
    assert np.array_equal(F, Fo)  # the promoted FP64 forward is the oracle's, bit for bit
    fd, unc = oracle_fd_bracket(cfg, nsteps, w0, w1)
    print("thermal", thermal, "F", F, "\nvjp", J, "\nfd ", fd)
    assert (np.abs(fd[0, :5]) > 1e-3 * np.abs(fd[0]).max()).sum() >= 3  # the yield law matters here
    assert_jacobian(J, fd, unc)


@pytest.mark.parametrize("thermal", ["adiabatic", "contact"])
def test_vjp_matches_host_chain(thermal):
    """The GPU reverse mode against the host restatement of the same chain (same blocks, pairs
    pushed instead of pulled, plain sums): equal to rounding."""
    from test_adjoint_chain import host_chain
    cfg = adj_config(thermal)
    nsteps, K, w0, w1 = 150, 25, 50, 150
    F, J, _ = gpu_jacobian(cfg, nsteps, K, w0, w1)
    for c, cot in enumerate(([1.0, 0.0], [0.0, 1.0])):
        Fh, gh, _, _ = host_chain(cfg, nsteps, w0, w1, cot)
        assert np.abs(F - Fh).max() <= 1e-12 * np.abs(F).max()
        err = np.abs(J[c] - gh) / np.abs(gh).max()
        print(thermal, c, "gpu", J[c], "\nhost", gh, "\nrel", err)
        assert (err < 1e-8).all()


def test_vjp_independent_of_checkpoint_interval():
    cfg = adj_config("contact")
    J = [gpu_jacobian(cfg, 60, K, 10, 60)[1] for K in (1, 7, 60)]
    assert np.array_equal(J[0], J[1]) and np.array_equal(J[0], J[2])


def test_vjp_leaves_the_run_state():
    cfg = adj_config("contact")
    _, _, sim = gpu_jacobian(cfg, 40, 8, 0, 40)
    ref = pymfc.Sim(cfg)
    ref.step(40)
    sim.step(10)
    ref.step(10)
    for f in ("pos", "vel", "S", "T", "eps_pl", "fc", "ft", "flags", "hash"):
        assert np.array_equal(sim.field(f), ref.field(f)), f
    assert np.array_equal(sim.force_history(), ref.force_history())


def test_autograd_calibration_loss():
    """torch.autograd.Function around mfc_vjp_forward/backward: the calibration loss
    |Fmean(theta) - F_target|^2 of a 100-step window, target from perturbed parameters."""
    torch = pytest.importorskip("torch")

    cfg = adj_config("adiabatic")
    nsteps, K, w0, w1 = 150, 30, 50, 150

    class MeanForce(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            sim = pymfc.Sim(with_theta(cfg, theta.detach().cpu().numpy()))
            F = sim.vjp_forward(nsteps, K, w0, w1)
            ctx.sim = sim
            return torch.tensor(F, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gF):
            return torch.tensor(ctx.sim.vjp_backward(gF.detach().cpu().numpy()), dtype=torch.float64)

    th0 = theta_of(cfg)
    rng = np.random.default_rng(7)
    th_target = th0 * (1 + rng.uniform(-0.05, 0.05, 6))
    th_target[5] = 0.30
    F_target = oracle_mean_force(with_theta(cfg, th_target), nsteps, w0, w1)

    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    loss = ((MeanForce.apply(theta) - torch.tensor(F_target)) ** 2).sum()
    loss.backward()
    g = theta.grad.numpy()

    def L(th):
        return float(((oracle_mean_force(with_theta(cfg, th), nsteps, w0, w1) - F_target) ** 2).sum())

    def fd_at(rel):
        out = np.zeros(6)
        for j in range(6):
            h = rel * abs(th0[j])
            tp, tm = th0.copy(), th0.copy()
            tp[j] += h
            tm[j] -= h
            out[j] = (L(tp) - L(tm)) / (2 * h)
        return out

    fd = fd_at(1e-6)
    unc = np.maximum(np.abs(fd - fd_at(1e-5)), np.abs(fd - fd_at(1e-7)))
    scale = np.abs(fd).max()
    err = np.abs(g - fd)
    print("loss", float(loss), "\ngrad", g, "\nfd  ", fd, "\nrel err", err / scale, "\nfd uncertainty", unc / scale)
    assert (err <= 1e-4 * scale + 2 * unc).all(), err / scale
