"""The oracle's neighbour sums against an O(N^2) brute force over all particle pairs.

The oracle visits only the 5 x 5 cells around a particle (interactions_gpu.cu:219-229); the
brute force takes every pair within the 2h support. On a jittered block with random
velocities, temperatures and densities (one step, so the predicted state is the initial state),
the continuity rate (-rho (v_x,x + v_y,y), actions_gpu.cu:164-174) and the Brookshaw heat rate
(interactions_gpu.cu:313-326, 341-343) must agree to rounding: a neighbour missed by the cell
window, a wrong cell id or a wrong cutoff shows up as an O(1) relative error. The pair
expressions are the cubic spline gradient of kernels.cuh:30-70 (float radius) evaluated
independently in numpy."""
import numpy as np

import pymfc
import oracle_binding as ob

PI_FAC = 10.0 / (7.0 * np.pi)


def grad_w(dx, dy, h):
    """cubic_spline gradient with the reference's float radius; zero outside (1e-8, 4h^2)."""
    rr2 = dx * dx + dy * dy
    inside = (rr2 < 4 * h * h) & (rr2 >= 1e-8)
    rad = np.sqrt(rr2.astype(np.float32)).astype(np.float64)
    rad = np.where(inside, rad, 1.0)
    h1 = 1.0 / h
    q = rad * h1
    m = 2.0 - q
    der = np.where(q > 1.0, -0.75 * m * m * h1 / rad, -3.0 * q * (1.0 - 0.75 * q) * h1 / rad)
    fac = PI_FAC * h1 * h1
    der = np.where(inside, der, 0.0)
    return der * dx * fac, der * dy * fac, inside, rr2


def block(thermal):
    c = pymfc.cut_config(16, 14, 10.0, thermal=thermal, jitter=0.2, seed=7, fixed_rows=0)
    c.fix_left = 0
    c.fix_right = 0
    c.sample_tool = 0
    c.tool_tl_x = -0.05  # tool far from the block: no contact
    return c


def test_continuity_rate_equals_all_pairs_sum():
    o = ob.OracleSim(block("adiabatic"))
    N = o.N
    rng = np.random.default_rng(3)
    v = rng.normal(0, 1e-4, (N, 2))
    rho = o.field("rho") * (1 + 0.01 * rng.random(N))
    o.set_field("vel", v)
    o.set_field("rho", rho)
    p, h = o.field("pos"), o.field("h")[0]
    mass = o.info.mass
    o.step(1)
    ids = o.field("id")
    rho_t = np.empty(N)
    rho_t[ids] = o.field("rho_t")
    dx = p[:, None, 0] - p[None, :, 0]
    dy = p[:, None, 1] - p[None, :, 1]
    wx, wy, inside, _ = grad_w(dx, dy, h)
    vol = mass / rho
    du = v[None, :, 0] - v[:, None, 0]
    dv = v[None, :, 1] - v[:, None, 1]
    vxx = (du * wx * vol[None, :]).sum(axis=1)
    vyy = (dv * wy * vol[None, :]).sum(axis=1)
    ref = -rho * (vxx + vyy)
    scale = np.abs(rho * (np.abs(du * wx * vol[None, :]).sum(axis=1) + np.abs(dv * wy * vol[None, :]).sum(axis=1)))
    assert inside.sum(axis=1).min() >= 5         # every particle has neighbours
    assert np.all(np.abs(rho_t - ref) <= 1e-12 * scale + 1e-300)


def test_brookshaw_rate_equals_all_pairs_sum():
    c = block("contact")
    o = ob.OracleSim(c)
    N = o.N
    rng = np.random.default_rng(4)
    T = 300.0 + 50.0 * rng.random(N)
    o.set_field("T", T)
    p, h, rho = o.field("pos"), o.field("h")[0], o.field("rho")
    mass = o.info.mass
    o.step(1)
    ids = o.field("id")
    T_t = np.empty(N)
    T_t[ids] = o.field("T_t")
    dx = p[:, None, 0] - p[None, :, 0]
    dy = p[:, None, 1] - p[None, :, 1]
    wx, wy, inside, rr2 = grad_w(dx, dy, h)
    rij = np.sqrt(rr2)
    ok = inside & (rij > 1e-8)
    rs = np.where(ok, rij, 1.0)
    e_dot_gw = (dx / rs) * wx + (dy / rs) * wy
    vol = mass / rho
    terms = np.where(ok, 2.0 * vol[None, :] * (T[:, None] - T[None, :]) / rs * e_dot_gw, 0.0)
    s = terms.sum(axis=1)
    # T_t = alpha * sum with alpha = k / (rho0 cp) of the workpiece; fit alpha on the largest
    # entries, then require every entry to match the all-pairs sum with that one constant
    big = np.abs(s) > 0.1 * np.abs(s).max()
    alpha = np.median(T_t[big] / s[big])
    assert alpha > 0
    scale = alpha * np.abs(terms).sum(axis=1)
    assert np.all(np.abs(T_t - alpha * s) <= 1e-10 * scale + 1e-300)
