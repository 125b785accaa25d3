"""VJPs of the adjoint building blocks (mfc_adjoint.h, built for the host by tests/harness)
against central differences of their forward companions.

Each block is checked on inputs at the magnitudes of the cutting benchmarks (positions in cm,
stresses in Mbar, velocities in cm/us), on every branch the step takes (in/out of the kernel
support, approaching/receding pairs, elastic/plastic, in/out of contact, stick/slip, fixed,
tool and blanked particles, PSE and Brookshaw conduction). A dropped term, a wrong sign or a
transposed index in a VJP shows up as a mismatch in the component it touches."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import pymfc

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "harness", "libadjh.so")

REC, LN, TS = 16, 14, 27
R_X, R_Y, R_FLAGS, R_H, R_U, R_V, R_RHO, R_VOL, R_AXX, R_AXY, R_AYY, R_RXX, R_RXY, R_RYY, R_C, R_T = range(16)
(TS_X, TS_Y, TS_U, TS_V, TS_SXX, TS_SXY, TS_SYY, TS_SZZ, TS_RHO, TS_T, TS_EPS, TS_EPSD, TS_H, TS_FCX, TS_FCY,
 TS_FTX, TS_FTY, TS_PTX, TS_PTY, TS_VTX, TS_VTY, TS_RHOT, TS_STXX, TS_STXY, TS_STYY, TS_STZZ, TS_TT) = range(27)
FIXED, BLANKED, TOOL = 1, 2, 4


def _lib():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.dirname(HARNESS)], check=True, capture_output=True)
    L = C.CDLL(HARNESS)
    vp, dp, i32, d = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double
    L.adjh_make_consts.argtypes = [C.POINTER(pymfc.Config), vp, C.POINTER(pymfc.ToolState)]
    L.adjh_set_theta.argtypes = [vp, dp]
    L.adjh_set_theta.restype = None
    L.adjh_set_secant_tol.argtypes = [vp, d]
    L.adjh_set_secant_tol.restype = None
    L.adjh_pair_fwd.argtypes = [vp, dp, dp, i32, i32, d, dp]
    L.adjh_pair_vjp.argtypes = [vp, dp, dp, i32, i32, d, dp, dp, dp]
    L.adjh_pse_fwd.argtypes = [vp, dp, dp]
    L.adjh_pse_fwd.restype = d
    L.adjh_pse_vjp.argtypes = [vp, dp, dp, d, dp, dp]
    L.adjh_art_fwd.argtypes = [dp, d, dp]
    L.adjh_art_vjp.argtypes = [dp, d, dp, dp]
    L.adjh_jc.argtypes = [vp, d, d, d, dp]
    L.adjh_contact.argtypes = [vp, C.POINTER(pymfc.ToolState), dp, dp, dp, dp]
    L.adjh_tail.argtypes = [vp, C.POINTER(pymfc.ToolState), dp, d, dp, i32, i32, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.adjh_record_fwd.argtypes = [vp, dp, i32, dp, dp, dp]
    L.adjh_record_vjp.argtypes = [vp, dp, i32, d, dp, dp]
    for f in ("adjh_pair_fwd", "adjh_pair_vjp", "adjh_pse_vjp", "adjh_art_fwd", "adjh_art_vjp", "adjh_jc",
              "adjh_tail", "adjh_record_fwd", "adjh_record_vjp"):
        getattr(L, f).restype = None
    return L


L = _lib()


def P(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Consts:
    """dev_consts of a configuration (promoted, tight secant) plus its initial tool."""

    def __init__(self, thermal="adiabatic", mu=0.35):
        cfg = pymfc.cut_config(40, 12, 10.0, thermal=thermal, mu=mu)
        cfg.promote_fp64 = 1
        cfg.secant_tol = 1e-15
        self.cfg = cfg
        self.buf = C.create_string_buffer(L.adjh_consts_size())
        self.tool = pymfc.ToolState()
        assert L.adjh_make_consts(C.byref(cfg), self.buf, C.byref(self.tool)) == 0
        self.theta = np.array([cfg.jc_A, cfg.jc_B, cfg.jc_n, cfg.jc_C, cfg.jc_m, cfg.mu])
        self.h = cfg.hdx * cfg.dx
        self.mass = cfg.dx * cfg.dx * cfg.rho0

    def with_theta(self, th: np.ndarray):
        b = C.create_string_buffer(self.buf.raw, len(self.buf.raw))
        L.adjh_set_theta(b, P(np.ascontiguousarray(th, dtype=np.float64)))
        return b


class FD(np.ndarray):
    """central differences with the rounding noise of each component attached (.noise)"""


def fd_grad(f, x: np.ndarray, scale: np.ndarray, rel: float = 1e-4) -> np.ndarray:
    """fourth-order central differences of the scalar f at x, step rel * scale per component
    (truncation error ~ rel^4); .noise holds each component's rounding bound 10 eps max|f| / h"""
    g = np.zeros_like(x)
    noise = np.zeros_like(x)
    for q in range(x.size):
        if scale[q] == 0:
            continue
        h = rel * scale[q]
        v = []
        for m in (2, 1, -1, -2):
            xx = x.copy()
            xx[q] += m * h
            v.append(f(xx))
        g[q] = (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)
        noise[q] = 10 * np.finfo(float).eps * max(abs(t) for t in v) / h
    out = g.view(FD)
    out.noise = noise
    return out


def assert_grad(g: np.ndarray, fd: np.ndarray, tol: float = 1e-6, what: str = ""):
    # componentwise: relative to the component's own size, plus the rounding noise of the
    # differences and a floor of 1e-9 of the largest component
    fd_v = np.asarray(fd)
    scale = np.maximum(np.abs(fd_v), np.abs(g))
    floor = 1e-9 * np.max(np.abs(fd_v)) if fd_v.size else 0.0
    noise = getattr(fd, "noise", np.zeros_like(fd_v))
    bad = np.abs(g - fd_v) > tol * scale + floor + noise
    assert not bad.any(), f"{what}: mismatch at {np.nonzero(bad)[0].tolist()}\nvjp={g[bad]}\nfd ={fd_v[bad]}"


# ------------------------------------------------------------------ records of a particle pair
def make_record(rng, k: Consts, x, y, T=None):
    rec = np.zeros(REC)
    rec[R_X], rec[R_Y], rec[R_H] = x, y, k.h
    rec[R_U], rec[R_V] = rng.normal(0, 2e-4, 2)
    rec[R_RHO] = k.cfg.rho0 * (1 + rng.uniform(-0.02, 0.02))
    rec[R_VOL] = k.mass / rec[R_RHO]
    rec[R_AXX:R_AYY + 1] = rng.normal(0, 3e-4, 3)
    rec[R_RXX:R_RYY + 1] = rng.normal(0, 2e-5, 3)
    rec[R_C] = np.sqrt(k.cfg.E / (3 * (1 - 2 * k.cfg.nu)) / rec[R_RHO])
    rec[R_T] = rng.uniform(300, 900) if T is None else T
    return rec


def rec_scale(rec: np.ndarray, h: float) -> np.ndarray:
    s = np.abs(rec).copy()
    s[R_X] = s[R_Y] = h
    s[R_FLAGS] = 0.0
    s[R_H] = 0.0  # h enters through the support radius only (uniform smoothing length)
    # the sums are linear in the stress records: a step well above their size keeps the
    # differences clear of rounding
    s[R_AXX:R_RYY + 1] = np.maximum(s[R_AXX:R_RYY + 1], 1e-2)
    s[R_VOL] = 1e-2
    return s


def pair_cases(rng, k: Consts):
    out = []
    for q in (0.35, 0.8, 1.3, 1.85):  # both spline pieces, away from the kink at q = 1
        for approach in (True, False):
            for ti, tj, ta in ((0, 0, 0.0), (0, 0, 1e-5), (1, 0, 1e-5), (0, 1, 1e-5), (1, 1, 1e-5)):
                ang = rng.uniform(0, 2 * np.pi)
                ri = make_record(rng, k, 0.02, 0.01)
                rj = make_record(rng, k, 0.02 + q * k.h * np.cos(ang), 0.01 + q * k.h * np.sin(ang))
                # relative velocity along -/+ the separation: approaching or receding pair
                e = np.array([ri[R_X] - rj[R_X], ri[R_Y] - rj[R_Y]])
                e /= np.linalg.norm(e)
                sgn = -1.0 if approach else 1.0
                rj[R_U], rj[R_V] = ri[R_U] - sgn * 3e-4 * e[0], ri[R_V] - sgn * 3e-4 * e[1]
                out.append((ri, rj, ti, tj, ta))
    return out


def test_pair_vjp_matches_differences():
    rng = np.random.default_rng(11)
    k = Consts()
    for ri, rj, ti, tj, ta in pair_cases(rng, k):
        lam = rng.normal(0, 1, LN)

        def J(x):
            o = np.zeros(LN)
            a, b = x[:REC].copy(), x[REC:].copy()
            L.adjh_pair_fwd(k.buf, P(a), P(b), ti, tj, ta, P(o))
            return float(lam @ o)

        gi, gj = np.zeros(REC), np.zeros(REC)
        L.adjh_pair_vjp(k.buf, P(ri), P(rj), ti, tj, ta, P(lam), P(gi), P(gj))
        x = np.concatenate([ri, rj])
        sc = np.concatenate([rec_scale(ri, k.h), rec_scale(rj, k.h)])
        fd = fd_grad(J, x, sc)
        assert_grad(np.concatenate([gi, gj]), fd, what=f"pair ti={ti} tj={tj} ta={ta}")


def test_pair_outside_support_is_zero():
    rng = np.random.default_rng(12)
    k = Consts()
    ri = make_record(rng, k, 0.02, 0.01)
    rj = make_record(rng, k, 0.02 + 2.05 * k.h, 0.01)
    o = np.zeros(LN)
    L.adjh_pair_fwd(k.buf, P(ri), P(rj), 0, 0, 1e-5, P(o))
    gi, gj = np.zeros(REC), np.zeros(REC)
    L.adjh_pair_vjp(k.buf, P(ri), P(rj), 0, 0, 1e-5, P(np.ones(LN)), P(gi), P(gj))
    assert not o.any() and not gi.any() and not gj.any()


def test_pse_vjp_matches_differences():
    rng = np.random.default_rng(13)
    k = Consts(thermal="pse")
    for q in (0.3, 1.1, 2.6, 4.5):  # PSE has no cutoff
        ri = make_record(rng, k, 0.02, 0.01)
        rj = make_record(rng, k, 0.02 + q * k.h * 0.6, 0.01 + q * k.h * 0.8)
        lam = rng.normal()

        def J(x):
            return lam * L.adjh_pse_fwd(k.buf, P(x[:REC].copy()), P(x[REC:].copy()))

        gi, gj = np.zeros(REC), np.zeros(REC)
        L.adjh_pse_vjp(k.buf, P(ri), P(rj), lam, P(gi), P(gj))
        x = np.concatenate([ri, rj])
        fd = fd_grad(J, x, np.concatenate([rec_scale(ri, k.h), rec_scale(rj, k.h)]))
        assert_grad(np.concatenate([gi, gj]), fd, what=f"pse q={q}")


def test_artificial_stress_vjp_matches_differences():
    rng = np.random.default_rng(14)
    eps = 0.3
    n_tension = 0
    for _ in range(40):
        rho = 4.43 * (1 + rng.uniform(-0.02, 0.02))
        x = np.array([rho, rng.normal(0, 5e-3), *rng.normal(0, 1e-2, 3)])
        bR = rng.normal(0, 1, 3)

        def J(v):
            R = np.zeros(3)
            L.adjh_art_fwd(P(v.copy()), eps, P(R))
            return float(bR @ R)

        R = np.zeros(3)
        L.adjh_art_fwd(P(x), eps, P(R))
        n_tension += bool(R.any())
        g = np.zeros(5)
        L.adjh_art_vjp(P(x), eps, P(bR), P(g))
        fd = fd_grad(J, x, np.array([rho, 1e-2, 1e-2, 1e-2, 1e-2]))
        assert_grad(g, fd, tol=1e-5, what="artificial stress")
    assert n_tension > 10  # the tension branch was exercised


def test_johnson_cook_partials():
    k = Consts()
    for e, ed, T in ((0.05, 0.3e-6, 400.0), (0.8, 5e-4, 900.0), (1.5, 2e-2, 1200.0)):
        out = np.zeros(9)
        L.adjh_jc(k.buf, e, ed, T, P(out))

        def s_of(v):
            o = np.zeros(9)
            L.adjh_jc(k.buf, v[0], v[1], v[2], P(o))
            return o[0]

        fd = fd_grad(s_of, np.array([e, ed, T]), np.array([e, ed, T]))
        assert_grad(out[1:4], fd, what="jc state partials")

        def s_theta(th):
            kb = k.with_theta(th)
            o = np.zeros(9)
            L.adjh_jc(kb, e, ed, T, P(o))
            return o[0]

        fdt = fd_grad(s_theta, k.theta.copy(), np.abs(k.theta))
        assert_grad(out[4:9], fdt[:5], what="jc parameter partials")


def contact_points(rng, k: Consts, n: int):
    """points inside the tool near its edge (rake face, fillet, clearance face)"""
    t = k.tool
    cx, cy = t.fillet_p[0], t.fillet_p[1]
    pts = []
    while len(pts) < n:
        x = cx + rng.uniform(-3, 3) * t.fillet_r
        y = cy + rng.uniform(-3, 3) * t.fillet_r
        inp = np.array([x, y, 0, 0, 0, 0.0])
        fwd, g = np.zeros(6), np.zeros(7)
        if L.adjh_contact(k.buf, C.byref(t), P(inp), P(np.zeros(6)), P(fwd), P(g)):
            pts.append((x, y))
    return pts


def test_contact_vjp_matches_differences():
    rng = np.random.default_rng(15)
    k = Consts()
    m, dt = k.mass, None
    checked = 0
    for (x, y) in contact_points(rng, k, 60):
        for slip in (False, True):
            u, v = rng.normal(0, 3e-4, 2)
            f0 = rng.normal(0, 1, 2) * (1e3 if slip else 1e-9)  # large history force -> slip
            inp = np.array([x, y, u, v, f0[0], f0[1]])
            bar = rng.normal(0, 1, 6)
            fwd, g = np.zeros(6), np.zeros(7)
            L.adjh_contact(k.buf, C.byref(k.tool), P(inp), P(bar), P(fwd), P(g))

            def F(vv, kb=k.buf):
                o, gg = np.zeros(6), np.zeros(7)
                L.adjh_contact(kb, C.byref(k.tool), P(vv[:6].copy()), P(np.zeros(6)), P(o), P(gg))
                return o

            # skip points within a step of a change of the closest primitive (normal jumps)
            hpos = 2e-4 * k.h
            jump = False
            for q in (0, 1):
                for s in (-1, 1):
                    xx = inp.copy()
                    xx[q] += s * hpos
                    if np.abs(F(xx)[4:] - fwd[4:]).max() > 1e-4:
                        jump = True
            if jump:
                continue
            def Jall(vv):
                th = k.theta.copy()
                th[5] = vv[6]
                return float(bar @ F(vv, k.with_theta(th)))

            x7 = np.concatenate([inp, [k.theta[5]]])
            sc = np.array([k.h, k.h, 3e-4, 3e-4, max(abs(f0[0]), 1e-4), max(abs(f0[1]), 1e-4), k.theta[5]])
            fd = fd_grad(Jall, x7, sc)
            assert_grad(g, fd, tol=1e-5, what=f"contact at {x},{y} slip={slip}")
            checked += 1
    assert checked > 60


# ------------------------------------------------------------------ tail of one particle
def make_state(rng, k: Consts, x, y, plastic: bool, ft0=(0.0, 0.0)):
    s = np.zeros(TS)
    s[TS_X], s[TS_Y] = x, y
    s[TS_U], s[TS_V] = rng.normal(0, 2e-4, 2)
    dev = 2.2e-2 if plastic else 2e-3
    s[TS_SXX], s[TS_SYY] = dev * (1 + 0.1 * rng.normal()), -dev * (1 + 0.1 * rng.normal())
    s[TS_SXY] = 0.3 * dev * rng.normal()
    s[TS_SZZ] = -(s[TS_SXX] + s[TS_SYY]) * 0.5
    s[TS_RHO] = k.cfg.rho0 * (1 + rng.uniform(-0.01, 0.01))
    s[TS_T] = rng.uniform(350, 700)
    s[TS_EPS] = rng.uniform(0.05, 0.5)
    s[TS_EPSD] = rng.uniform(1e-4, 1e-2)
    s[TS_H] = k.h
    s[TS_FCX], s[TS_FCY] = rng.normal(0, 1e-6, 2)
    s[TS_FTX], s[TS_FTY] = ft0
    s[TS_PTX:TS_VTY + 1] = rng.normal(0, 1e-4, 4)
    s[TS_RHOT] = rng.normal(0, 1e-3)
    s[TS_STXX:TS_STZZ + 1] = rng.normal(0, 1e-3, 4)
    s[TS_TT] = rng.normal(0, 1.0)
    return s


def state_scale(s: np.ndarray, k: Consts) -> np.ndarray:
    sc = np.maximum(np.abs(s), 1e-12)
    sc[TS_X] = sc[TS_Y] = k.h
    sc[TS_H] = 0.0
    return sc


def run_tail(k: Consts, kb, s, p, n_prev, flags, quirk, acc, out_bar, nbar, Fbar):
    g_s, g_n, g_acc, g_th, fwd = np.zeros(TS), np.zeros(2), np.zeros(LN), np.zeros(6), np.zeros(TS + 2)
    L.adjh_tail(kb, C.byref(k.tool), P(s), p, P(n_prev), flags, quirk, P(acc), P(out_bar), P(nbar), P(Fbar),
                P(g_s), P(g_n), P(g_acc), P(g_th), P(fwd))
    return g_s, g_n, g_acc, g_th, fwd


def check_tail(rng, k: Consts, s, flags, acc, n_prev, quirk=0, what=""):
    p = 1e-3
    out_bar = rng.normal(0, 1, TS)
    nbar = rng.normal(0, 1, 2)
    Fbar = rng.normal(0, 1, 2)

    def J(x):
        th = x[TS + 2 + LN:]
        kb = k.with_theta(th)
        _, _, _, _, fwd = run_tail(k, kb, x[:TS].copy(), p, x[TS:TS + 2].copy(), flags, quirk,
                                   x[TS + 2:TS + 2 + LN].copy(), out_bar, nbar, Fbar)
        F = np.array([fwd[TS_FCX] + fwd[TS_FTX], fwd[TS_FCY] + fwd[TS_FTY]])
        return float(out_bar @ fwd[:TS] + nbar @ fwd[TS:] + Fbar @ F)

    kb = k.with_theta(k.theta)
    g_s, g_n, g_acc, g_th, fwd = run_tail(k, kb, s, p, n_prev, flags, quirk, acc, out_bar, nbar, Fbar)
    x = np.concatenate([s, n_prev, acc, k.theta])
    sc = np.concatenate([state_scale(s, k), np.maximum(np.abs(n_prev), 1e-3), np.maximum(np.abs(acc), 1e-12),
                         np.abs(k.theta)])
    fd = fd_grad(J, x, sc)
    assert_grad(np.concatenate([g_s, g_n, g_acc, g_th]), fd, tol=2e-5, what=what)
    return fwd


def make_acc(rng):
    acc = np.zeros(LN)
    acc[0:4] = rng.normal(0, 2e-3, 4)      # velocity gradient, 1/us
    acc[4:8] = rng.normal(0, 1e-4, 4)      # stress divergence, cm/us^2
    acc[8:12] = rng.normal(0, 1e-5, 4)     # viscous and XSPH terms
    acc[12:14] = rng.normal(0, 1e6, 2)     # Laplacian sums (Brookshaw, PSE), K/cm^2
    return acc


@pytest.mark.parametrize("thermal", ["adiabatic", "contact", "pse"])
def test_tail_vjp_free_particles(thermal):
    rng = np.random.default_rng(16)
    k = Consts(thermal=thermal)
    n_plastic = 0
    for case in range(12):
        plastic = case % 2 == 1
        flags = [0, 0, FIXED, TOOL, BLANKED, 0][case % 6]
        s = make_state(rng, k, 0.03 + 0.001 * case, 0.01, plastic, ft0=tuple(rng.normal(0, 1e-5, 2)))
        fwd = check_tail(rng, k, s, flags, make_acc(rng), rng.normal(0, 1, 2), quirk=case % 3 == 0,
                         what=f"{thermal} case {case} flags {flags}")
        n_plastic += (flags & TOOL) == 0 and fwd[TS_EPSD] > 0
    assert n_plastic >= 3


def test_tail_vjp_in_contact():
    rng = np.random.default_rng(17)
    k = Consts(thermal="contact")
    dt = None
    n = 0
    for (x, y) in contact_points(rng, k, 10):
        for plastic in (False, True):
            s = make_state(rng, k, x, y, plastic, ft0=tuple(rng.normal(0, 1e-5, 2)))
            s[TS_PTX:TS_PTY + 1] = 0.0  # corrector leaves the particle where it is (inside the tool)
            check_tail(rng, k, s, 0, make_acc(rng), rng.normal(0, 1, 2), what=f"contact {x},{y} plastic={plastic}")
            n += 1
    assert n == 20


# ------------------------------------------------------------------ record
@pytest.mark.parametrize("flags", [0, TOOL, BLANKED])
def test_record_vjp_matches_differences(flags):
    rng = np.random.default_rng(18)
    k = Consts()
    for case in range(8):
        s = make_state(rng, k, 0.02, 0.01, plastic=case % 2 == 1)
        p0, R0 = 2e-3, np.array([1e-5, -2e-5, 3e-6])
        brec = rng.normal(0, 1, REC)
        brec[R_FLAGS] = brec[R_H] = 0.0

        def J(x):
            p = np.array([p0])
            R = R0.copy()
            rec = np.zeros(REC)
            L.adjh_record_fwd(k.buf, P(x.copy()), flags, P(p), P(R), P(rec))
            return float(brec @ rec)

        p = np.array([p0])
        R = R0.copy()
        rec = np.zeros(REC)
        L.adjh_record_fwd(k.buf, P(s), flags, P(p), P(R), P(rec))
        gs = np.zeros(TS)
        L.adjh_record_vjp(k.buf, P(s), flags, float(p[0]), P(brec), P(gs))
        fd = fd_grad(J, s, state_scale(s, k))
        assert_grad(gs, fd, tol=1e-5, what=f"record flags {flags} case {case}")


def test_tail_vjp_secant_fold_branch():
    """Plastic states whose return equation g(dl) = |S| - 2 G dl - sqrt(2/3) sigma_Y = 0 has no
    root: the norm counts S_xy once while the yield test uses the von Mises stress, so a
    shear-dominated stress between 0.577 and 0.816 of the yield stress is plastic with g(0) < 0.
    The secant then stops at the fixed point of its fold branch, which the VJP differentiates."""
    rng = np.random.default_rng(19)
    k = Consts()
    n_fold = 0
    for case in range(8):
        s = make_state(rng, k, 0.03, 0.01, plastic=False)
        s[TS_STXX:TS_STZZ + 1] = 0.0
        jc = np.zeros(9)
        L.adjh_jc(k.buf, s[TS_EPS], 0.0, s[TS_T], P(jc))  # rate-free yield stress
        s[TS_SXX] = s[TS_SYY] = s[TS_SZZ] = 0.0
        s[TS_SXY] = 0.72 * jc[0]
        acc = make_acc(rng)
        acc[0:4] *= 1e-3  # keep the stress rate small against the stress
        fwd = check_tail(rng, k, s, 0, acc, rng.normal(0, 1, 2), what=f"fold case {case}")
        g0 = np.sqrt(2.0 / 3.0) * jc[0] - 0.72 * jc[0]
        n_fold += int(fwd[TS_EPS] > s[TS_EPS] and g0 > 0)
    assert n_fold >= 6
