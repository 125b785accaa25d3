"""Pins the oracle's per-pair and per-particle functions with closed forms, quadrature,
finite differences, eigen-decompositions and special cases (the reference ships no tests or
golden vectors). Each check is chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails it."""
import math

import numpy as np
import pytest

import oracle_binding as ob

H = 1.5e-3  # smoothing length of the 10 um lattice, cm


# ----------------------------------------------------------------------------- cubic spline
def W(r, h=H):
    return ob.cubic_spline(r, 0.0, 0.0, 0.0, h)[0]


def _textbook_w(r, h=H):
    q = r / h
    val = np.where(q <= 1, 1 - 1.5 * q * q * (1 - 0.5 * q), np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
    return 10 / (7 * np.pi * h * h) * val


def test_cubic_spline_normalisation_by_quadrature():
    """integral of W over the plane = 1 for the M4 spline: the oracle's q <= 1 piece with the
    textbook tail integrates to 1 (pins PI_FAC = 10/(7 pi) and the inner polynomial)."""
    r = np.linspace(0.0, 2 * H, 40001)
    w = np.array([W(x) if (x * x >= 1e-8 and x <= H) else 0.0 for x in r])
    w = np.where((r * r < 1e-8) | (r > H), _textbook_w(r), w)
    inner = r <= H
    # agreement to the float rounding of the radius (rad = sqrtf(rr2))
    assert np.allclose(w[inner & (r * r >= 1e-8)], _textbook_w(r[inner & (r * r >= 1e-8)]), rtol=5e-7)
    integral = np.trapezoid(w * 2 * np.pi * r, r)
    assert abs(integral - 1.0) < 1e-6


def test_cubic_spline_outer_value_is_sixth_power():
    """For 1 < q < 2 the reference's value is 0.25 (2-q)^6 (kernels.cuh:54-57 squares (2-q)
    and multiplies three squares); its gradient is the textbook -0.75 (2-q)^2 / h."""
    for q in [1.1, 1.4, 1.7, 1.95]:
        r = q * H
        q32 = float(np.sqrt(np.float32(r * r))) / H
        want = 0.454728408833987 / H / H * 0.25 * (2 - q32) ** 6
        assert abs(W(r) - want) < 1e-12 * want
        g = ob.cubic_spline(r, 0.0, 0.0, 0.0, H)
        rad = float(np.sqrt(np.float32(r * r)))
        assert abs(g[1] - (-0.75 * (2 - q32) ** 2 / H / rad) * r * 0.454728408833987 / H / H) < 1e-10 * abs(g[1])


def test_cubic_spline_support_and_cutoffs():
    assert W(2 * H) == 0.0 and W(2 * H * 1.0000001) == 0.0
    assert W(0.99e-4) == 0.0  # rr2 < 1e-8 cm^2: pairs closer than 1 um are dropped (kernels.cuh:43)
    assert W(1.01e-4) > 0.0
    # C1 at q = 1: value and slope continuous
    a, b = W(H * (1 - 1e-9)), W(H * (1 + 1e-9))
    assert abs(a - b) / a < 1e-7


def test_cubic_spline_gradient_is_gradient_wrt_xi():
    """w.y, w.z = dW/dx_i, dW/dy_i of the M4 spline (interactions use x_i - x_j), by central
    differences of the textbook kernel (the oracle's own value differs for q > 1, see above)."""
    rng = np.random.default_rng(0)
    for _ in range(300):
        r = rng.uniform(0.2 * H, 1.95 * H)
        th = rng.uniform(0, 2 * np.pi)
        xi, yi = r * np.cos(th), r * np.sin(th)
        g = ob.cubic_spline(xi + 0.01, yi + 0.02, 0.01, 0.02, H)
        d = 1e-9
        fx = (_textbook_w(math.hypot(xi + d, yi)) - _textbook_w(math.hypot(xi - d, yi))) / (2 * d)
        fy = (_textbook_w(math.hypot(xi, yi + d)) - _textbook_w(math.hypot(xi, yi - d))) / (2 * d)
        scale = max(abs(fx), abs(fy), 1e-3 * _textbook_w(0.0) / H)
        assert abs(g[1] - fx) < 2e-4 * scale and abs(g[2] - fy) < 2e-4 * scale


def test_cubic_spline_uses_float_radius():
    """rad = sqrtf(rr2) (kernels.cuh:50): the value follows the float-rounded radius."""
    r2 = 1.234567891234e-6
    rad32 = float(np.sqrt(np.float32(r2)))
    rad64 = math.sqrt(r2)
    assert rad32 != rad64
    q = rad32 / H
    der32 = -3.0 * q * (1 - 0.75 * q) * (1 / H) / rad32
    g = ob.cubic_spline(math.sqrt(r2), 0.0, 0.0, 0.0, H)
    fac = 0.454728408833987 * (1 / H) * (1 / H)
    assert abs(g[1] - der32 * math.sqrt(r2) * fac) <= 1e-15 * abs(g[1])


# ----------------------------------------------------------------------------- PSE kernel
def test_pse_kernel_second_moment_gives_laplacian():
    """sum (T_j - T_i) eta_ij V_j -> lap T: for T = x^2 the plane integral of x^2 eta is 2."""
    h = 1e-3
    x = np.linspace(-6 * h, 6 * h, 601)
    X, Y = np.meshgrid(x, x)
    dA = (x[1] - x[0]) ** 2
    eta = np.vectorize(lambda a, b: ob.lapl_pse(a, b, 0.0, 0.0, h))(X, Y)
    assert abs(np.sum(X * X * eta) * dA - 2.0) < 1e-3
    assert abs(np.sum(eta) * dA * h * h / 4 - 1.0) < 1e-3  # 4/(pi h^4) e^{-r^2/h^2}: mass 4/h^2


# ----------------------------------------------------------------------------- Johnson-Cook
JC = np.array([0.01098, 0.01092, 0.014, 0.93, 1.1, 1836.0, 300.0, 1e-6])  # A B C n m Tmelt Tref eps_dot_ref


def test_sigma_yield_special_points():
    A, B, Cc, n, m, Tm, Tr, e0 = JC
    assert ob.sigma_yield(JC, 0.0, 0.0, Tr) == A                        # pow(0, n) = 0, (1+0)^C = 1
    assert abs(ob.sigma_yield(JC, 1.0, 0.0, Tr) - (A + B)) < 1e-17      # pow(1, n) = 1
    assert ob.sigma_yield(JC, 0.3, 0.0, Tm) == 0.0                      # theta = 1
    half = ob.sigma_yield(JC, 0.0, 0.0, Tr + 0.5 * (Tm - Tr))
    assert abs(half - A * (1 - 0.5 ** m)) < 1e-16
    # rate term: (1 + eps*)^C at eps* <= 1, 1 + C ln eps* above (plasticity.cuh:38-42)
    assert abs(ob.sigma_yield(JC, 0.0, e0, Tr) - A * 2.0 ** Cc) < 1e-16
    assert abs(ob.sigma_yield(JC, 0.0, 2 * e0, Tr) - A * (1 + Cc * math.log(2.0))) < 1e-16
    assert abs(ob.sigma_yield(JC, 0.0, 1e3 * e0, Tr) - A * (1 + Cc * math.log(1e3))) < 1e-16
    # below the reference temperature theta < 0 and pow(theta, m) is NaN, as in the reference
    assert math.isnan(ob.sigma_yield(JC, 0.1, 0.0, Tr - 1.0))


def test_sigma_yield_monotone():
    e = np.linspace(0.0, 2.0, 50)
    s = [ob.sigma_yield(JC, x, 0.0, 400.0) for x in e]
    assert np.all(np.diff(s) > 0)
    T = np.linspace(300.0, 1800.0, 50)
    s = [ob.sigma_yield(JC, 0.2, 0.0, t) for t in T]
    assert np.all(np.diff(s) < 0)


def _bisect(f, lo, hi):
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return lo


def test_secant_root_matches_bisection():
    """With a tight tolerance the secant update (perturbation x1.01, reflection of negative
    iterates, plasticity.cuh:94-118) converges to the root of g found by bisection; with the
    reference's 1e-6 it stops within 101 iterations."""
    G, dt = 0.424, 8e-4
    rng = np.random.default_rng(1)
    for _ in range(100):
        eps0 = rng.uniform(0, 1.0)
        T = rng.uniform(300, 1200)
        epsd = 10 ** rng.uniform(-6, -2)
        sy = ob.sigma_yield(JC, eps0, epsd, T)
        norm = math.sqrt(2 / 3) * sy * rng.uniform(1.0005, 1.5)
        init = max(epsd * dt * math.sqrt(2 / 3), 1e-8)
        root = _bisect(lambda x: ob.opt_fun(JC, x, norm, eps0, T, dt, G), 0.0, norm / (2 * G))
        dl, it = ob.solve_secant(JC, init, 1e-15, norm, eps0, T, dt, G)
        assert abs(dl - root) < 1e-9 * root, (dl, root, it)
        dl, it = ob.solve_secant(JC, init, 1e-6, norm, eps0, T, dt, G)
        assert it <= 101 and 0 <= dl


def test_secant_absolute_tolerance_stops_early():
    """The stop test is |dl - dl_old| < 1e-6 in absolute terms (plasticity.cuh:108): from the
    floor 1e-8 with a large overstress the first update is already below it."""
    G, dt = 0.424, 8e-4
    sy = ob.sigma_yield(JC, 0.5, 0.0, 600.0)
    norm = math.sqrt(2 / 3) * sy * 2.0
    dl, it = ob.solve_secant(JC, 1e-8, 1e-6, norm, 0.5, 600.0, dt, G)
    root = _bisect(lambda x: ob.opt_fun(JC, x, norm, 0.5, 600.0, dt, G), 0.0, norm / (2 * G))
    assert it == 0 and dl < 1e-6 < root
    dl2, it2 = ob.solve_secant(JC, 1e-8, 1e-12, norm, 0.5, 600.0, dt, G)
    assert abs(dl2 - root) < 1e-9 * root and it2 > 0


def test_radial_return_consistency():
    """Plastic: S scaled radially, ||S_new|| = ||S_trial|| - 2 G dl, eps_pl grows by sqrt(2/3) dl,
    heating tq/(cp rho) d_eps sigma_Y; elastic: state kept and eps_pl_dot reset."""
    G, dt, tq, cp = 0.424, 8e-4, 0.9, 553e-8
    S = np.array([0.02, 0.004, -0.015, -0.005])
    p = 0.0
    state = np.array([*S, p, 0.05, 1e-4, 350.0, 4.43])
    out = ob.radial_return(JC, G, dt, tq, cp, state)
    norm = float(np.sqrt(np.float32(S @ S)))
    dl_f = (out[5] - 0.05) / float(np.sqrt(np.float32(2 / 3)))
    ratio = out[:4] / S
    assert np.allclose(ratio, ratio[0], rtol=1e-12) and 0 < ratio[0] < 1
    dl = (1 - ratio[0]) * norm / (2 * G)
    assert abs(dl - dl_f) < 1e-7 * dl
    assert out[6] == (out[5] - 0.05) / dt or abs(out[6] - (out[5] - 0.05) / dt) < 1e-12 * out[6]
    sy = ob.sigma_yield(JC, out[5], out[6], 350.0)
    assert abs((out[7] - 350.0) - tq / (cp * 4.43) * (out[5] - 0.05) * sy) < 1e-9
    # elastic: tiny stress
    st2 = np.array([1e-4, 0.0, -1e-4, 0.0, 0.0, 0.05, 1e-4, 350.0, 4.43])
    out2 = ob.radial_return(JC, G, dt, tq, cp, st2)
    assert np.array_equal(out2[:6], st2[:6]) and out2[6] == 0.0 and out2[7] == 350.0


# ----------------------------------------------------------------------------- artificial stress
def test_artificial_stress_against_eigendecomposition():
    rng = np.random.default_rng(2)
    eps, rho = 0.3, 4.43
    for _ in range(300):
        S = rng.normal(0, 0.01, 4)
        p = rng.normal(0, 0.01)
        R = ob.artificial_stress(rho, p, S, eps)
        sig = np.array([[S[0] - p, S[1]], [S[1], S[2] - p]])
        lam, V = np.linalg.eigh(sig)
        Rp = np.where(lam > 0, -eps * lam / rho ** 2, 0.0)
        Rm = V @ np.diag(Rp) @ V.T
        scale = max(np.abs(Rm).max(), 1e-3 * eps * np.abs(lam).max() / rho ** 2)
        err = np.abs(np.array([R[0] - Rm[0, 0], R[1] - Rm[0, 1], R[2] - Rm[1, 1]])).max()
        assert err < 2e-6 * scale, (S, p, R, Rm)


# ----------------------------------------------------------------------------- Jaumann rate
def test_jaumann_rate_matrix_form_and_deviatoric():
    rng = np.random.default_rng(3)
    G = 0.424
    for _ in range(200):
        vd = rng.normal(0, 1.0, 4)  # vx,x vx,y vy,x vy,y
        S = rng.normal(0, 0.01, 4)
        out = ob.jaumann(G, vd, S)
        L = np.array([[vd[0], vd[1], 0], [vd[2], vd[3], 0], [0, 0, 0]])  # L_ab = dv_a/dx_b
        D = 0.5 * (L + L.T)
        Om = 0.5 * (L.T - L)  # the reference's omega: omega[col 0][row 1] = (vy,x - vx,y)/2
        Sm = np.array([[S[0], S[1], 0], [S[1], S[2], 0], [0, 0, S[3]]])
        St = 2 * G * (D - np.trace(D) / 3 * np.eye(3)) + Om.T @ Sm + Sm @ Om
        want = np.array([St[0, 0], St[0, 1], St[1, 1], St[2, 2]])
        assert np.allclose(out, want, rtol=1e-13, atol=1e-16 * np.abs(want).max())
        # the spin terms are trace-free and the strain-rate part deviatoric
        assert abs(out[0] + out[2] + out[3]) < 1e-13 * np.abs(out).max()


def test_jaumann_rigid_spin_rotates_stress():
    """For a pure spin the rate equals d/dt (Q S Q^T) at t = 0 (objectivity)."""
    G, w = 0.424, 0.7
    vd = np.array([0.0, -w, w, 0.0])  # v = w (-y, x)
    S = np.array([0.01, 0.003, -0.004, 0.001])
    out = ob.jaumann(G, vd, S)
    Sm = np.array([[S[0], S[1]], [S[1], S[2]]])
    dt = 1e-7
    Q = np.array([[np.cos(w * dt), -np.sin(w * dt)], [np.sin(w * dt), np.cos(w * dt)]])
    rate = (Q @ Sm @ Q.T - Q.T @ Sm @ Q) / (2 * dt)
    assert np.allclose([out[0], out[1], out[2]], [rate[0, 0], rate[0, 1], rate[1, 1]], rtol=1e-6)
    assert out[3] == 0.0
