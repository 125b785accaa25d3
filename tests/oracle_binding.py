"""ctypes binding of the FP64 host oracle (oracle/liboracle.so) for the tests and the
benchmark's CPU-baseline leg. Same interface shape as pymfc.Sim."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import pymfc
from pymfc import Config, Info, ToolState, FIELDS, INT_FIELDS

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MFC_ORACLE_LIB selects another build of the oracle (the mutation tests load deliberately
# broken copies through it)
ORACLE_PATH = os.environ.get("MFC_ORACLE_LIB") or os.path.join(REPO, "oracle", "liboracle.so")

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run make -C oracle")
        L = C.CDLL(ORACLE_PATH)
        vp, i32, i64, dp, sz = C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double), C.c_size_t
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_config_ref_cut.argtypes = [C.POINTER(Config), i32] + [C.c_double] * 5
        L.oracle_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_destroy.restype = None
        L.oracle_step.argtypes = [vp, i32]
        L.oracle_get_info.argtypes = [vp, C.POINTER(Info)]
        L.oracle_tool_force.argtypes = [vp, dp]
        L.oracle_force_history.argtypes = [vp, dp, i64, C.POINTER(C.c_int64)]
        L.oracle_get_field.argtypes = [vp, i32, vp, sz]
        L.oracle_set_field.argtypes = [vp, i32, vp, sz]
        L.oracle_get_cells.argtypes = [vp, C.POINTER(C.c_int32), sz]
        L.oracle_get_tool.argtypes = [vp, C.POINTER(ToolState)]
        L.oracle_tool_wear.argtypes = [vp, C.c_double, C.c_double, dp, dp]
        L.oracle_plastic_iterations.argtypes = [vp]
        L.oracle_plastic_iterations.restype = C.c_int64
        L.oracle_out_of_grid.argtypes = [vp]
        L.oracle_out_of_grid.restype = C.c_int64
        L.oracle_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [i32]
        L.oracle_set_threads.restype = None
        L.oracle_cubic_spline.argtypes = [C.c_double] * 5 + [dp]
        L.oracle_cubic_spline.restype = None
        L.oracle_lapl_pse.argtypes = [C.c_double] * 5
        L.oracle_lapl_pse.restype = C.c_double
        L.oracle_sigma_yield.argtypes = [dp, C.c_double, C.c_double, C.c_double]
        L.oracle_sigma_yield.restype = C.c_double
        L.oracle_sigma_yield_tanh.argtypes = [dp, C.c_double, C.c_double, C.c_double]
        L.oracle_sigma_yield_tanh.restype = C.c_double
        L.oracle_opt_fun.argtypes = [dp] + [C.c_double] * 6
        L.oracle_opt_fun.restype = C.c_double
        L.oracle_solve_secant.argtypes = [dp] + [C.c_double] * 7 + [C.POINTER(C.c_int)]
        L.oracle_solve_secant.restype = C.c_double
        L.oracle_radial_return.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.oracle_radial_return.restype = None
        L.oracle_artificial_stress.argtypes = [C.c_double, C.c_double, dp, C.c_double, dp]
        L.oracle_artificial_stress.restype = None
        L.oracle_jaumann.argtypes = [C.c_double, dp, dp, dp]
        L.oracle_jaumann.restype = None
        L.oracle_contact_point.argtypes = [vp] + [C.c_double] * 6 + [dp]
        L.oracle_contact_point.restype = None
        L.oracle_math.argtypes = [i32, dp, dp, dp, i64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().oracle_last_error().decode()}")


def config_ref_cut(ny=30, rake=10.0, clear=8.0, chamfer=0.002, speed=83.333328e-5 / 500.0 * 70.0, feed=0.01) -> Config:
    c = Config()
    _check(lib().oracle_config_ref_cut(C.byref(c), ny, rake, clear, chamfer, speed, feed), "oracle_config_ref_cut")
    return c


class OracleSim:
    def __init__(self, cfg: Config):
        self._h = C.c_void_p()
        _check(lib().oracle_create(C.byref(cfg), C.byref(self._h)), "oracle_create")
        self.cfg = cfg.copy()

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self):
        i = Info()
        _check(lib().oracle_get_info(self._h, C.byref(i)), "oracle_get_info")
        return pymfc.SimInfo(*[getattr(i, n) for n, _ in Info._fields_])

    @property
    def N(self):
        return self.info.n_particles

    def step(self, n=1):
        _check(lib().oracle_step(self._h, int(n)), "oracle_step")

    def tool_force(self):
        F = np.zeros(2)
        _check(lib().oracle_tool_force(self._h, _p(F)), "oracle_tool_force")
        return F

    def tool_wear(self, K=0.078, alpha=2500.0, tool_vel=(0.0, 0.0)):
        """(min, max, mean, count) of the Usui wear rate over the particles in contact."""
        out = np.zeros(4)
        tv = np.array(tool_vel, dtype=np.float64)
        _check(lib().oracle_tool_wear(self._h, K, alpha, _p(tv), _p(out)), "oracle_tool_wear")
        return out

    def force_history(self):
        n = C.c_int64()
        _check(lib().oracle_force_history(self._h, None, 0, C.byref(n)), "oracle_force_history")
        buf = np.zeros((max(n.value, 1), 2))
        _check(lib().oracle_force_history(self._h, _p(buf), n.value, C.byref(n)), "oracle_force_history")
        return buf[: n.value].copy()

    def field(self, name):
        fid, nc = FIELDS[name]
        N = self.N
        dt = np.int32 if name in INT_FIELDS else np.float64
        out = np.zeros((N, nc) if nc > 1 else (N,), dtype=dt)
        _check(lib().oracle_get_field(self._h, fid, out.ctypes.data_as(C.c_void_p), out.nbytes), f"oracle_get_field({name})")
        return out

    def set_field(self, name, values):
        fid, nc = FIELDS[name]
        dt = np.int32 if name in INT_FIELDS else np.float64
        v = np.ascontiguousarray(values, dtype=dt)
        _check(lib().oracle_set_field(self._h, fid, v.ctypes.data_as(C.c_void_p), v.nbytes), f"oracle_set_field({name})")

    def cells(self):
        nc = self.info.num_cell
        out = np.zeros(nc + 1, dtype=np.int32)
        _check(lib().oracle_get_cells(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), nc + 1), "oracle_get_cells")
        return out

    def tool(self):
        t = ToolState()
        _check(lib().oracle_get_tool(self._h, C.byref(t)), "oracle_get_tool")
        return t.as_arrays()

    def plastic_iterations(self):
        return lib().oracle_plastic_iterations(self._h)

    def out_of_grid(self):
        return lib().oracle_out_of_grid(self._h)

    def contact_point(self, x, y, vx=0.0, vy=0.0, ftx=0.0, fty=0.0):
        out = np.zeros(9)
        lib().oracle_contact_point(self._h, x, y, vx, vy, ftx, fty, _p(out))
        return out


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def threads() -> int:
    return lib().oracle_threads()


def host_math(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    yp = None if y is None else _p(np.ascontiguousarray(y, dtype=np.float64))
    if y is not None:
        y = np.ascontiguousarray(y, dtype=np.float64)
        yp = _p(y)
    _check(lib().oracle_math(fn, _p(x), yp, _p(out), x.size), "oracle_math")
    return out


def cubic_spline(xi, yi, xj, yj, h):
    out = np.zeros(3)
    lib().oracle_cubic_spline(xi, yi, xj, yj, h, _p(out))
    return out


def lapl_pse(xi, yi, xj, yj, h):
    return lib().oracle_lapl_pse(xi, yi, xj, yj, h)


def _jc(jc):
    return np.ascontiguousarray(jc, dtype=np.float64)


def sigma_yield(jc, eps, epsd, T):
    a = _jc(jc)
    return lib().oracle_sigma_yield(_p(a), eps, epsd, T)


def sigma_yield_tanh(jc, tanh, eps, epsd, T):
    """jc as for sigma_yield, tanh = (a, b, c, d)."""
    a = np.concatenate([_jc(jc), np.asarray(tanh, dtype=np.float64)])
    return lib().oracle_sigma_yield_tanh(_p(a), eps, epsd, T)


def opt_fun(jc, dl, norm, eps, T, dt, G):
    a = _jc(jc)
    return lib().oracle_opt_fun(_p(a), dl, norm, eps, T, dt, G)


def solve_secant(jc, init, tol, norm, eps, T, dt, G):
    a = _jc(jc)
    it = C.c_int()
    r = lib().oracle_solve_secant(_p(a), init, tol, norm, eps, T, dt, G, C.byref(it))
    return r, it.value


def radial_return(jc, G, dt, tq, cp, state):
    a = _jc(jc)
    s = np.ascontiguousarray(state, dtype=np.float64).copy()
    lib().oracle_radial_return(_p(a), G, dt, tq, cp, _p(s))
    return s


def artificial_stress(rho, p, S, eps):
    Sa = np.ascontiguousarray(S, dtype=np.float64)
    out = np.zeros(3)
    lib().oracle_artificial_stress(rho, p, _p(Sa), eps, _p(out))
    return out


def jaumann(G, vder, S):
    v = np.ascontiguousarray(vder, dtype=np.float64)
    s = np.ascontiguousarray(S, dtype=np.float64)
    out = np.zeros(4)
    lib().oracle_jaumann(G, _p(v), _p(s), _p(out))
    return out
