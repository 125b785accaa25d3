"""Closed-form pins of the oracle's per-particle actions and pair terms, each checked through the
oracle's whole step on small explicit configurations (one isolated particle, one pair, or an
interior of a lattice), against values that the mathematics fixes rather than a second
restatement of the same code:

- equation of state (actions_gpu.cu:41-51): p = c0^2 (rho - rho0), c0 = sqrtf(K / rho0) in
  single precision, at the half-step density rho0 + dt/2 rho_t;
- leap-frog predictor and corrector (leap_frog.cu:42-92): an isolated particle with constant
  derivatives moves exactly by x1 = x0 + dt (v0 + dt/2 a), and its rates are the predicted ones;
- the temperature predictor's half step: the next Brookshaw rate is the Laplacian of
  T0 + dt/2 T_t (interactions_gpu.cu:313-326);
- the stress divergence of a linear compressive stress field, entering at its half-step value
  (interactions_gpu.cu:268-280, actions_gpu.cu:196-210);
- two-particle closed forms: Monaghan artificial viscosity pushes an approaching pair apart with
  Pi = (-alpha c mu + beta mu^2) / rho (:282-305); XSPH moves a particle by
  -eps W m / rho_ij v_ij (:308-309); the artificial-stress term enters as f^4 (R_i + R_j) with
  f = W(r) / W(dp), checked at two separations (:268-280);
- friction heating of one particle by hand, with the reference's scalar subtraction in the
  relative velocity (actions_gpu.cu:135-162);
- the fixed-particle boundary conditions (actions_gpu.cu:296-311) and the thermal one (:287-294).
"""
import numpy as np
import pytest

import pymfc
import oracle_binding as ob

PI_FAC = 10.0 / (7.0 * np.pi)


def small_config(nx, ny, thermal="adiabatic"):
    """nx x ny lattice at dx = 10 um, nothing fixed, the tool far to the left (never in
    contact), no tool particles."""
    c = pymfc.cut_config(nx, ny, 10.0, thermal=thermal, fixed_rows=0)
    c.fix_left = 0
    c.fix_right = 0
    c.sample_tool = 0
    c.tool_tl_x = -0.05
    return c


def moduli(cfg):
    G = cfg.E / (2.0 * (1.0 + cfg.nu))
    K = 2.0 * G * (1.0 + cfg.nu) / (3.0 * (1.0 - 2.0 * cfg.nu))
    return G, K


def c_f32(K, rho):
    """sqrtf of the reference: the square root of the single-precision argument, rounded to
    single precision (IEEE-754 square root is correctly rounded)."""
    return float(np.sqrt(np.float32(K / rho)))


def spline(dxy, h):
    """Cubic spline of kernels.cuh:30-70 in closed form: W and its gradient with respect to
    the first particle for the separation dxy = x_i - x_j (outer piece as the reference
    writes it, 0.25 (2 - q)^6, with the derivative of 0.25 (2 - q)^3)."""
    r = float(np.hypot(dxy[0], dxy[1]))
    q = r / h
    fac = PI_FAC / (h * h)
    if q >= 2.0 or r * r < 1e-8:
        return 0.0, np.zeros(2)
    if q > 1.0:
        val = 0.25 * (2.0 - q) ** 6
        der = -0.75 * (2.0 - q) ** 2 / (h * r)
    else:
        val = 1.0 - 1.5 * q * q * (1.0 - 0.5 * q)
        der = -3.0 * q * (1.0 - 0.75 * q) / (h * r)
    return val * fac, der * fac * np.asarray(dxy, dtype=np.float64)


def by_id(o, name):
    """Field values indexed by particle id."""
    v = o.field(name)
    out = np.empty_like(v)
    out[o.field("id")] = v
    return out


def interior(pos, h, margin):
    m = margin * h
    x, y = pos[:, 0], pos[:, 1]
    return (x > x.min() + m) & (x < x.max() - m) & (y > y.min() + m) & (y < y.max() - m)


# ------------------------------------------------------------------ equation of state
def test_eos_at_the_half_step_density():
    cfg = small_config(24, 16)
    o = ob.OracleSim(cfg)
    rng = np.random.default_rng(11)
    N = o.N
    rho = cfg.rho0 * (1.0 + rng.uniform(-2e-3, 2e-3, N))
    rho_t = rng.uniform(-1.0, 1.0, N) * 1e-3
    o.set_field("rho", rho)
    o.set_field("rho_t", rho_t)
    ids0 = o.field("id")
    dt = o.info.dt
    o.step(1)
    _, K = moduli(cfg)
    c0 = c_f32(K, cfg.rho0)
    rho_half = np.empty(N)
    rho_half[ids0] = rho + 0.5 * dt * rho_t
    want = c0 * c0 * (rho_half - cfg.rho0)
    got = by_id(o, "p")  # written at the particle's sorted position of this step
    assert np.allclose(got, want, rtol=1e-13, atol=0.0), np.max(np.abs(got / want - 1))
    # c0 (not c0^2) or the full-step density would be off by orders of magnitude / by dt rho_t
    assert np.max(np.abs(got - want)) < 1e-3 * np.max(np.abs(0.5 * dt * rho_t * c0 * c0))


# ------------------------------------------------------------------ leap-frog integrator
def test_isolated_particle_moves_by_the_constant_derivative_formulas():
    cfg = small_config(1, 1)
    o = ob.OracleSim(cfg)
    dt = o.info.dt
    x0 = o.field("pos")[0].copy()
    v0 = np.array([3e-5, -2e-5])
    a = np.array([4e-3, 7e-3])        # vel_t of the previous step
    b = np.array([1e-4, -5e-5])       # pos_t of the previous step (only moves the predicted point)
    S0 = np.array([[-1e-3, 2e-4, -5e-4, 3e-4]])
    St = np.array([[1e-2, -3e-3, 2e-3, 1e-3]])
    o.set_field("vel", v0[None, :])
    o.set_field("vel_t", a[None, :])
    o.set_field("pos_t", b[None, :])
    o.set_field("S", S0)
    o.set_field("S_t", St)
    o.set_field("T_t", np.array([0.25]))
    T0 = o.field("T")[0]
    o.step(1)
    vhalf = v0 + 0.5 * dt * a
    # no neighbour, no contact: pos_t = predicted velocity, vel_t = 0, Jaumann rate of a
    # particle without velocity gradient = 0, so v and S keep their start values
    assert np.array_equal(o.field("pos_t")[0], vhalf)
    assert np.array_equal(o.field("vel_t")[0], [0.0, 0.0])
    assert np.array_equal(o.field("pos")[0], x0 + dt * vhalf)
    assert np.array_equal(o.field("vel")[0], v0)
    assert np.array_equal(o.field("S")[0], S0[0])
    assert np.array_equal(o.field("S_t")[0], [0.0, 0.0, 0.0, 0.0])
    # adiabatic: T_t is never rewritten (the reference writes it only with conduction), so the
    # corrector moves T by the full step of the old rate
    assert o.field("T")[0] == T0 + dt * 0.25
    # a second step from the new state: constant velocity, straight line
    x1 = o.field("pos")[0].copy()
    o.step(1)
    assert np.array_equal(o.field("pos")[0], x1 + dt * v0)


def test_temperature_predictor_takes_half_a_step():
    """T0 uniform, previous rate c x^2: the predicted field T0 + dt/2 c x^2 has the Laplacian
    dt c, so the next Brookshaw rate is alpha dt c in the interior (a full-step predictor
    would give twice that)."""
    cfg = small_config(30, 24, thermal="contact")
    o = ob.OracleSim(cfg)
    p = o.field("pos")
    dt = o.info.dt
    c = 1e7
    o.set_field("T_t", c * (p[:, 0] - p[:, 0].min()) ** 2)
    inner = interior(p, o.field("h")[0], 3.5)
    ids = o.field("id")
    o.step(1)
    alpha = cfg.k_wp / (cfg.rho0 * cfg.cp)
    k = np.isin(o.field("id"), ids[inner])
    Tt = o.field("T_t")[k]
    assert np.allclose(Tt, alpha * dt * c, rtol=0.05), (Tt.min(), Tt.max(), alpha * dt * c)


# ------------------------------------------------------------------ stress divergence
def test_linear_compressive_stress_divergence():
    """sigma_xx(x) linear and compressive at the half step (S0 = 0, S_t = -g (x - x_min + 5 dx)):
    no artificial stress (no tensile principal stress), p = 0, so vel_t = (1/rho) d sigma/dx
    = -dt/2 g / rho inside; a full-step stress predictor would double it."""
    cfg = small_config(30, 24)
    o = ob.OracleSim(cfg)
    p = o.field("pos")
    dt = o.info.dt
    g = 2e2
    St = np.zeros((o.N, 4))
    St[:, 0] = -g * (p[:, 0] - p[:, 0].min() + 5 * cfg.dx)
    o.set_field("S_t", St)
    inner = interior(p, o.field("h")[0], 2.2)
    ids = o.field("id")
    o.step(1)
    k = np.isin(o.field("id"), ids[inner])
    vt = o.field("vel_t")[k]
    want = -0.5 * dt * g / cfg.rho0
    assert np.allclose(vt[:, 0], want, rtol=0.03), (vt[:, 0].min(), vt[:, 0].max(), want)
    assert np.all(np.abs(vt[:, 1]) < 1e-6 * abs(want))


# ------------------------------------------------------------------ two-particle closed forms
def pair(r_over_dx, thermal="adiabatic"):
    cfg = small_config(2, 1, thermal=thermal)
    o = ob.OracleSim(cfg)
    p = o.field("pos")
    ids = o.field("id")
    x0 = p[ids == 0][0]
    pos = np.empty((2, 2))
    pos[ids == 0] = x0
    pos[ids == 1] = x0 + [r_over_dx * cfg.dx, 0.0]
    o.set_field("pos", pos)
    return cfg, o


def set_by_id(o, name, values):
    ids = o.field("id")
    o.set_field(name, np.asarray(values)[ids])


def test_viscosity_pushes_an_approaching_pair_apart():
    cfg, o = pair(1.2)
    u = 2e-3
    set_by_id(o, "vel", [[u, 0.0], [-u, 0.0]])
    x = by_id(o, "pos")
    dt = o.info.dt
    # the previous step's position rates move the pair's predicted points closer by
    # dt b = 0.1 dx (half a step each of +-b); the velocities' rates are zero
    b = 0.1 * cfg.dx / dt
    set_by_id(o, "pos_t", [[b, 0.0], [-b, 0.0]])
    h = 1.5 * cfg.dx
    o.step(1)
    G, K = moduli(cfg)
    mass = cfg.dx * cfg.dx * cfg.rho0
    xp = x + 0.5 * dt * np.array([[b, 0.0], [-b, 0.0]])
    dxy = xp[0] - xp[1]
    W, gW = spline(dxy, h)
    vij = np.array([2 * u, 0.0])
    vdotx = vij @ dxy
    assert vdotx < 0
    c = c_f32(K, cfg.rho0)
    mu = h * vdotx / (dxy @ dxy + cfg.av_eta ** 2 * h * h)
    Pi = (-cfg.av_alpha * c * mu + cfg.av_beta * mu * mu) / cfg.rho0
    vt = by_id(o, "vel_t")
    want = -mass * Pi * gW
    assert np.allclose(vt[0], want, rtol=1e-6, atol=1e-30 * abs(want[0]))
    assert vt[0, 0] < 0 < vt[1, 0]                      # pushed apart
    assert np.allclose(vt[1], -vt[0], rtol=1e-12)      # equal and opposite
    # XSPH: pos_t = v - eps W m / rho_ij (v_i - v_j)
    pt = by_id(o, "pos_t")
    xs = -cfg.xsph_eps * W * mass / cfg.rho0 * vij
    assert np.allclose(pt[0], [u, 0.0] + xs, rtol=1e-6)
    assert np.allclose(pt[1], [-u, 0.0] - xs, rtol=1e-6)
    # a receding pair feels no viscosity
    cfg, o = pair(1.2)
    set_by_id(o, "vel", [[-u, 0.0], [u, 0.0]])
    o.step(1)
    assert np.array_equal(by_id(o, "vel_t"), np.zeros((2, 2)))


@pytest.mark.parametrize("r_over_dx", [1.3, 1.7])
def test_artificial_stress_enters_with_the_fourth_power_of_w(r_over_dx):
    """A pair under uniaxial tension s along the line joining them (p = 0, rho = rho0): each
    particle's artificial stress is R_xx = -eps s / rho^2 (its tensile principal stress), and the
    momentum rate is vel_t,x = 2 m s (1 - eps f^4) dW/dx / rho^2 with f = W(r) / W(dx)."""
    cfg, o = pair(r_over_dx)
    s = 2e-3
    S = np.zeros((2, 4))
    S[:, 0] = s
    o.set_field("S", S)
    x = by_id(o, "pos")
    h = 1.5 * cfg.dx
    o.step(1)
    mass = cfg.dx * cfg.dx * cfg.rho0
    W, gW = spline(x[0] - x[1], h)
    Wdp, _ = spline(np.array([cfg.dx, 0.0]), h)
    f = W / Wdp
    R = by_id(o, "R")
    assert np.allclose(R[:, 0], -cfg.stress_eps * s / cfg.rho0 ** 2, rtol=1e-12)
    assert np.all(R[:, 1:] == 0.0)
    vt = by_id(o, "vel_t")
    want = 2.0 * mass * s * (1.0 - cfg.stress_eps * f ** 4) * gW[0] / cfg.rho0 ** 2
    assert abs(vt[0, 0] / want - 1.0) < 1e-6, (vt[0, 0], want, f)
    assert vt[0, 1] == 0.0
    # the f-dependence is not linear: f^4 differs from f by far more than the tolerance here
    assert abs(f ** 4 - f) * cfg.stress_eps > 1e-2


# ------------------------------------------------------------------ friction heat
def test_friction_heat_of_one_particle_by_hand():
    """T += eta dt |ft| |v_r| / (cp m) with the corrected velocity, the previous step's friction
    force and contact normal, and v_r = v_d - (v_d . n) subtracted from both components (the
    reference's form, actions_gpu.cu:155-157; the tangential projection would differ)."""
    cfg = small_config(1, 1)
    o = ob.OracleSim(cfg)
    dt = o.info.dt
    mass = cfg.dx * cfg.dx * cfg.rho0
    ft = np.array([-3e-9, 4e-9])
    n = np.array([0.6, 0.8])
    v0 = np.array([1e-4, 2e-4])
    o.set_field("ft", ft[None, :])
    o.set_field("n", n[None, :])
    o.set_field("vel", v0[None, :])
    T0 = o.field("T")[0]
    o.step(1)
    v1 = v0 + dt * (ft / mass)           # momentum: vel_t = ft / m for a particle alone
    assert np.allclose(o.field("vel")[0], v1, rtol=1e-15)
    vtool = np.array(o.tool()["vel"])
    vd = v1 - vtool
    vr = vd - (vd @ n)
    f32 = lambda a: float(np.sqrt(np.float32(a)))
    want = cfg.eta * dt * f32(ft @ ft) * f32(vr @ vr) / (cfg.cp * mass)
    T1 = o.field("T")[0]
    assert abs(T1 - (T0 + want)) <= 4 * np.spacing(T0), (T1 - T0, want)
    proj = vd - (vd @ n) * n               # the projection the physics would use
    assert abs(np.linalg.norm(proj) / np.linalg.norm(vr) - 1.0) > 0.05
    # the contact pass clears the friction force of a particle away from the tool
    assert np.array_equal(o.field("ft")[0], [0.0, 0.0])


# ------------------------------------------------------------------ boundary conditions
def test_fixed_particles_hold_zero_velocity_force_and_rates():
    cfg = pymfc.cut_config(30, 12, 10.0, thermal="contact")
    o = ob.OracleSim(cfg)
    rng = np.random.default_rng(3)
    o.set_field("vel", rng.normal(0.0, 1e-4, (o.N, 2)))
    o.set_field("T", 300.0 + rng.uniform(0.0, 50.0, o.N))
    for _ in range(3):
        o.step(4)
        fl = o.field("flags")
        fixed = (fl & pymfc.FLAG_FIXED) != 0
        tool = (fl & pymfc.FLAG_TOOL) != 0
        wp = fixed & ~tool
        assert wp.sum() > 50 and (fixed & tool).sum() > 10
        for f in ("vel", "fc", "pos_t", "vel_t"):
            assert np.all(o.field(f)[wp] == 0.0), f
        # thermal boundary condition: every fixed particle, tool ones included, at T_init
        assert np.all(o.field("T")[fixed] == cfg.T_init)
        assert np.any(o.field("T")[~fixed] != cfg.T_init)
