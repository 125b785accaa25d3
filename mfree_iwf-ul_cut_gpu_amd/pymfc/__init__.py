"""ctypes binding of the MI355X SPH cutting solver (C ABI in include/mfc.h).

The Python side mirrors the reference's host interface for the cutting step:

    reference (src/)                                   here
    setup_ref_cut(ny, &grid, rake, clear, ...)          config_ref_cut(...) + Sim(cfg)
        benchmarks.cpp:255-537
    leap_frog::step(particles, grid)                    Sim.step(n)
        leap_frog.cu:94-155
    report_force(particles)  output.cu:31-39            Sim.tool_force()
    cudaMemcpy of particle_gpu fields                   Sim.field(name)

Loading fails loudly when lib/libmfc.so is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
import time
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmfc.so")

MFC_OK = 0
THERMAL_BROOKSHAW, THERMAL_PSE = 0, 1
PREC_FP64, PREC_FP32C = 0, 1
JC_PLAIN, JC_TANH = 0, 1  # MFC_JC_*: the flow stress of the radial return's secant
LATTICE_REF, LATTICE_EXPLICIT, LATTICE_RINGS, LATTICE_IMPACT = 0, 1, 2, 3
GRID_GREEN, GRID_ROTHLIN = 0, 1  # MFC_GRID_*
FLAG_FIXED, FLAG_BLANKED, FLAG_TOOL = 1, 2, 4

FIELDS = {
    "pos": (0, 2), "vel": (1, 2), "S": (2, 4), "rho": (3, 1), "p": (4, 1), "T": (5, 1),
    "eps_pl": (6, 1), "eps_pl_dot": (7, 1), "fc": (8, 2), "ft": (9, 2), "n": (10, 2), "h": (11, 1),
    "flags": (12, 1), "id": (13, 1), "hash": (14, 1), "pos_t": (15, 2), "vel_t": (16, 2),
    "rho_t": (17, 1), "S_t": (18, 4), "T_t": (19, 1), "R": (20, 3),
}
INT_FIELDS = {"flags", "id", "hash"}


class Config(C.Structure):
    """mfc_config (include/mfc.h)."""
    _fields_ = [
        ("lattice", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double),
        ("lo_x", C.c_double), ("hi_x", C.c_double), ("hi_y", C.c_double), ("hdx", C.c_double),
        ("jitter", C.c_double), ("seed", C.c_uint64),
        ("fixed_bottom_rows", C.c_int32), ("fix_left", C.c_int32), ("fix_right", C.c_int32),
        ("E", C.c_double), ("nu", C.c_double), ("rho0", C.c_double),
        ("jc_A", C.c_double), ("jc_B", C.c_double), ("jc_C", C.c_double), ("jc_n", C.c_double),
        ("jc_m", C.c_double), ("jc_Tmelt", C.c_double), ("jc_Tref", C.c_double), ("jc_eps_dot_ref", C.c_double),
        ("cp", C.c_double), ("tq", C.c_double), ("k_wp", C.c_double), ("eta", C.c_double), ("T_init", C.c_double),
        ("av_alpha", C.c_double), ("av_beta", C.c_double), ("av_eta", C.c_double), ("xsph_eps", C.c_double),
        ("stress_eps", C.c_double),
        ("tool_tl_x", C.c_double), ("tool_tl_y", C.c_double), ("tool_length", C.c_double), ("tool_height", C.c_double),
        ("rake_deg", C.c_double), ("clear_deg", C.c_double), ("edge_radius", C.c_double),
        ("feed", C.c_double), ("speed", C.c_double), ("mu", C.c_double), ("contact_alpha", C.c_double),
        ("sample_tool", C.c_int32), ("rho_tool", C.c_double), ("cp_tool", C.c_double), ("k_tool", C.c_double),
        ("thermal_solver", C.c_int32), ("precision", C.c_int32), ("cfl", C.c_double),
        ("pse_flag_quirk", C.c_int32), ("promote_fp64", C.c_int32), ("secant_tol", C.c_double),
        ("jc_model", C.c_int32), ("elastic", C.c_int32),
        ("jc_tanh_a", C.c_double), ("jc_tanh_b", C.c_double), ("jc_tanh_c", C.c_double), ("jc_tanh_d", C.c_double),
        ("grid_adaptive", C.c_int32), ("max_cell_factor", C.c_int32), ("grid_kind", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]

    def copy(self) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        return c

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if not n.startswith("reserved")}


class Info(C.Structure):
    """mfc_info (include/mfc.h)."""
    _fields_ = [
        ("n_particles", C.c_int32), ("n_tool", C.c_int32), ("grid_nx", C.c_int32), ("grid_ny", C.c_int32),
        ("num_cell", C.c_int32), ("precision", C.c_int32), ("dt", C.c_double), ("t_final", C.c_double),
        ("mass", C.c_double), ("bbmin_x", C.c_double), ("bbmin_y", C.c_double), ("cell_size", C.c_double),
        ("step", C.c_int64), ("out_of_grid", C.c_int64), ("grid_overflow", C.c_int64),
        ("grid_adaptive", C.c_int32), ("cell_table", C.c_int32),
    ]


class ToolState(C.Structure):
    """mfc_tool_state (include/mfc.h)."""
    _fields_ = [
        ("left", (C.c_double * 2) * 5), ("right", (C.c_double * 2) * 5), ("n", (C.c_double * 2) * 5),
        ("line_a", C.c_double * 5), ("line_b", C.c_double * 5), ("vertical", C.c_int32 * 5),
        ("num_seg", C.c_int32), ("has_fillet", C.c_int32),
        ("fillet_p", C.c_double * 2), ("fillet_r", C.c_double), ("fillet_t1", C.c_double),
        ("fillet_t2", C.c_double), ("vel", C.c_double * 2),
    ]

    def as_arrays(self) -> dict:
        ns = self.num_seg
        return {
            "left": np.array(self.left)[:ns], "right": np.array(self.right)[:ns], "n": np.array(self.n)[:ns],
            "line_a": np.array(self.line_a)[:ns], "line_b": np.array(self.line_b)[:ns],
            "vertical": np.array(self.vertical)[:ns], "has_fillet": self.has_fillet,
            "fillet": np.array([self.fillet_p[0], self.fillet_p[1], self.fillet_r, self.fillet_t1, self.fillet_t2]),
            "vel": np.array(self.vel),
        }


class VjpReq(C.Structure):
    """mfc_vjp_req (include/mfc.h)."""
    _fields_ = [("nsteps", C.c_int32), ("checkpoint_every", C.c_int32), ("w0", C.c_int32), ("w1", C.c_int32),
                ("cot", C.c_double * 2), ("Fmean", C.c_double * 2)]


STAGE_CAP = 2048  # MFC_STAGE_CAP: window candidates a tile can stage in LDS
XCD_CHUNK_MAX = 64  # MFC_XCD_CHUNK_MAX: longest XCD chunk of the interaction kernels' tile map
SORT_FAST_MAX = 1536  # MFC_SORT_FAST_MAX: most particles of a chunk the sort holds in registers
GRAPH_STEPS = 8   # steps per captured hipGraph of a handle (one graph per buffer parity)


# the launch configurations the autotuner tries: fused or split interaction, direct or compacted
# pair loop, and the LDS stage limit (0: tiles stage up to STAGE_CAP candidates; n > 0: tiles
# whose candidates span more than n sorted indices read them from memory; 1: no tile stages)
TUNING_SPACE = {"split_interact": (0, 1), "compact_pairs": (0, 1), "stage_limit": (0, 1536, 1), "xcd_chunk": (8, 32)}


def autotune(handle, steps: int = 100, space: dict | None = None, rounds: int = 2) -> dict:
    """Times every combination of the launch configuration space on the live handle (a Sim or a
    Batch; every configuration gives the same results, the run just advances) over `steps`
    steps each, and keeps the fastest. Each candidate first runs 2 * GRAPH_STEPS untimed steps,
    so that the hipGraph of the handle's buffer parity (unchanged by an even number of steps) is
    captured outside the timed window, and the
    candidates are timed in `rounds` interleaved passes (the run's contact state evolves under
    them), each keeping its minimum. The handle is left on the chosen settings with its graphs
    built. Returns the chosen settings, the ms per step of each candidate and the number of
    steps the handle advanced (steps_run)."""
    space = TUNING_SPACE if space is None else space
    keys = list(space)
    combos = list(itertools.product(*(space[k] for k in keys)))
    cand = {c: float("inf") for c in combos}
    ran = 0
    for _ in range(max(1, rounds)):
        for combo in combos:
            handle.set_tuning(**dict(zip(keys, combo)))
            handle.step(2 * GRAPH_STEPS)
            handle.sync()
            t0 = time.perf_counter()
            handle.step(steps)
            handle.sync()
            cand[combo] = min(cand[combo], (time.perf_counter() - t0) / steps * 1e3)
            ran += 2 * GRAPH_STEPS + steps
    best = min(cand, key=cand.get)
    handle.set_tuning(**dict(zip(keys, best)))
    handle.step(2 * GRAPH_STEPS)
    handle.sync()
    ran += 2 * GRAPH_STEPS
    label = {"split_interact": "split", "compact_pairs": "compact", "stage_limit": "stage", "xcd_chunk": "chunk"}
    out = dict(zip(keys, best))
    out["steps_run"] = ran
    out["ms_per_step"] = {"_".join(f"{label.get(k, k)}{v}" for k, v in zip(keys, c)): ms for c, ms in cand.items()}
    return out


class Tuning(C.Structure):
    """mfc_tuning (include/mfc.h)."""
    _fields_ = [("split_interact", C.c_int32), ("compact_pairs", C.c_int32), ("stage_limit", C.c_int32),
                ("xcd_chunk", C.c_int32), ("sort_fast", C.c_int32), ("sort_list", C.c_int32),
                ("reserved", C.c_int32 * 2)]


THETA = ("jc_A", "jc_B", "jc_n", "jc_C", "jc_m", "mu")  # order of the adjoint's parameters

_lib = None


def lib() -> C.CDLL:
    """The solver library; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dp, sz = C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double), C.c_size_t
        L.mfc_last_error.restype = C.c_char_p
        L.mfc_config_ref_cut.argtypes = [C.POINTER(Config), i32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.mfc_config_rings.argtypes = [C.POINTER(Config), i32]
        L.mfc_config_impact.argtypes = [C.POINTER(Config), i32]
        L.mfc_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.mfc_destroy.argtypes = [vp]
        L.mfc_destroy.restype = None
        L.mfc_step.argtypes = [vp, i32]
        L.mfc_sync.argtypes = [vp]
        L.mfc_profile.argtypes = [vp, i32, dp]
        L.mfc_tool_force.argtypes = [vp, dp]
        L.mfc_force_history.argtypes = [vp, dp, i64, C.POINTER(C.c_int64)]
        L.mfc_force_history_reset.argtypes = [vp]
        L.mfc_get_info.argtypes = [vp, C.POINTER(Info)]
        L.mfc_get_field.argtypes = [vp, i32, vp, sz]
        L.mfc_set_field.argtypes = [vp, i32, vp, sz]
        L.mfc_get_cells.argtypes = [vp, C.POINTER(C.c_int32), sz]
        L.mfc_get_tool.argtypes = [vp, C.POINTER(ToolState)]
        L.mfc_tool_wear.argtypes = [vp, C.c_double, C.c_double, dp, C.c_double, dp]
        L.mfc_tool_wear_noalloc.argtypes = [vp, C.c_double, C.c_double, dp, C.c_double, dp]
        L.mfc_write_vtk.argtypes = [vp, C.c_char_p]
        L.mfc_write_tool_vtk.argtypes = [vp, C.c_char_p]
        L.mfc_device_math.argtypes = [i32, dp, dp, dp, i64]
        L.mfc_scenario.argtypes = [C.POINTER(Config), C.POINTER(Info), dp, dp, dp, dp, C.POINTER(C.c_int32),
                                   C.POINTER(ToolState)]
        L.mfc_field_components.argtypes = [i32]
        L.mfc_vjp_forward.argtypes = [vp, i32, i32, i32, i32, dp]
        L.mfc_vjp_backward.argtypes = [vp, dp, dp]
        L.mfc_vjp.argtypes = [vp, C.POINTER(VjpReq), dp]
        L.mfc_batch_create.argtypes = [C.POINTER(Config), i32, C.POINTER(vp)]
        L.mfc_batch_destroy.argtypes = [vp]
        L.mfc_batch_destroy.restype = None
        L.mfc_batch_size.argtypes = [vp]
        L.mfc_batch_instance.argtypes = [vp, i32, C.POINTER(vp)]
        L.mfc_batch_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.mfc_batch_step.argtypes = [vp, i32]
        L.mfc_batch_sync.argtypes = [vp]
        L.mfc_batch_vjp_forward.argtypes = [vp, i32, i32, i32, i32, dp]
        L.mfc_batch_vjp_backward.argtypes = [vp, dp, dp]
        L.mfc_save.argtypes = [vp, C.c_char_p]
        L.mfc_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mfc_get_config.argtypes = [vp, C.POINTER(Config)]
        L.mfc_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.mfc_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        _lib = L
    return _lib


class MfcError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != MFC_OK:
        raise MfcError(f"{what} failed ({rc}): {lib().mfc_last_error().decode()}")


def config_ref_cut(ny: int = 30, rake: float = 10.0, clear: float = 8.0, chamfer: float = 0.002,
                   speed: float = 83.333328e-5 / 500.0 * 70.0, feed: float = 0.01) -> Config:
    """setup_ref_cut's constants; defaults are main()'s (mfree_gpu.cpp:160-165)."""
    c = Config()
    _check(lib().mfc_config_ref_cut(C.byref(c), ny, rake, clear, chamfer, speed, feed), "mfc_config_ref_cut")
    return c


def config_rings(nbox: int = 80, precision: int = PREC_FP64) -> Config:
    """setup_rings (benchmarks.cpp:155-253): two rubber-like rings colliding at 2 x 200, elastic,
    no tool; nbox = 80 is main()'s commented call (mfree_gpu.cpp:209)."""
    c = Config()
    _check(lib().mfc_config_rings(C.byref(c), nbox), "mfc_config_rings")
    c.precision = precision
    return c


def config_impact(nbox: int = 20, precision: int = PREC_FP64, adaptive: bool = True) -> Config:
    """setup_impact (benchmarks.cpp:23-153): a steel 4340 disk (SI units) at 1e3 into a wall,
    Johnson-Cook plasticity without heat, no tool; nbox = 20 is main()'s commented call. The
    grid adapts to the particles every step as the reference's active line does
    (benchmarks.cpp:135); adaptive=False selects its commented fixed grid (:136)."""
    c = Config()
    _check(lib().mfc_config_impact(C.byref(c), nbox), "mfc_config_impact")
    c.precision = precision
    if not adaptive:
        c.grid_adaptive = 0
    return c


# ---------------------------------------------------------------- SI -> bomb units (cm, us, g, Mbar)
def um(x): return x * 1e-4                   # micrometre -> cm
def m_per_min(v): return v / 60.0 * 1e-4     # m/min -> cm/us
def gpa(x): return x * 1e-2                  # GPa -> Mbar
def mpa(x): return x * 1e-5                  # MPa -> Mbar
def kg_m3(x): return x * 1e-3                # kg/m^3 -> g/cm^3
def j_kgk(x): return x * 1e-8                # J/(kg K) -> Mbar cm^3/(g K)
def w_mk(x): return x * 1e-13                # W/(m K)
def per_s(x): return x * 1e-6                # 1/s -> 1/us


def cut_config(nx: int, ny: int, dx_um: float, *, thermal: str = "adiabatic", speed_m_min: float = 70.0,
               feed_um: float = 100.0, jitter: float = 0.0, seed: int = 0x355, precision: int = PREC_FP64,
               cfl: float = 0.3, k_wp: float = 7.1, k_tool: float = 50.0, mu: float = 0.35,
               jc: tuple | None = None, fixed_rows: int = 3) -> Config:
    """An orthogonal cut of an nx x ny Ti6Al4V lattice (Johnson-Cook after Johnson 1985 in SI
    units converted to the solver's units) with a 0 deg rake / 7 deg clearance tool of 5 um edge
    radius, placed like setup_ref_cut places its tool (benchmarks.cpp:389-418).

    thermal: "adiabatic" (no conduction, no tool particles), "contact" (Brookshaw conduction in
    the workpiece and into sampled tool particles), or "pse" (PSE conduction, no tool particles).
    jc: (A MPa, B MPa, n, C, m) overrides the Johnson-Cook constants."""
    c = config_ref_cut(ny=30, rake=0.0, clear=7.0, chamfer=um(5.0), speed=m_per_min(speed_m_min), feed=um(feed_um))
    c.lattice = LATTICE_EXPLICIT
    c.nx, c.ny, c.dx = nx, ny, um(dx_um)
    c.lo_x, c.hi_y = 0.005, 0.060
    # front of the tool 10 um beyond the workpiece edge, so that the placement rule of
    # setup_ref_cut (benchmarks.cpp:400-405) moves its rake face back onto the edge: the cut
    # starts at step 0 instead of after the ~1 mm approach of the reference's placement
    c.tool_tl_x = c.lo_x + um(10.0) - c.tool_length
    c.jitter, c.seed = jitter, seed
    c.fixed_bottom_rows = fixed_rows
    c.E, c.nu, c.rho0 = gpa(113.8), 0.342, kg_m3(4430.0)
    A, B, n, Cc, m = jc if jc is not None else (1098.0, 1092.0, 0.93, 0.014, 1.1)
    c.jc_A, c.jc_B, c.jc_n, c.jc_C, c.jc_m = mpa(A), mpa(B), n, Cc, m
    c.jc_Tref, c.jc_Tmelt, c.jc_eps_dot_ref = 300.0, 1836.0, per_s(1.0)
    c.cp, c.tq, c.eta, c.T_init = j_kgk(553.0), 0.9, 0.9, 300.0
    c.mu = mu
    c.cfl = cfl
    c.precision = precision
    if thermal == "adiabatic":
        c.k_wp, c.k_tool, c.sample_tool, c.thermal_solver = 0.0, 0.0, 0, THERMAL_BROOKSHAW
    elif thermal == "contact":
        c.k_wp, c.k_tool, c.sample_tool, c.thermal_solver = w_mk(k_wp), w_mk(k_tool), 1, THERMAL_BROOKSHAW
    elif thermal == "pse":
        c.k_wp, c.k_tool, c.sample_tool, c.thermal_solver = w_mk(k_wp), 0.0, 0, THERMAL_PSE
    else:
        raise ValueError(thermal)
    return c


# named workloads (sizes and physics of the benchmark cases)
def workload(name: str, precision: int = PREC_FP64) -> Config:
    if name == "adiabatic_6k":        # 150 x 40 at 10 um, FP64 adiabatic, 2000 steps
        return cut_config(150, 40, 10.0, thermal="adiabatic", precision=precision)
    if name == "thermal_contact_150k":  # 750 x 200 at 2 um, Brookshaw + tool thermal contact, 20000 steps
        return cut_config(750, 200, 2.0, thermal="contact", precision=precision)
    if name == "rings_250k":          # setup_rings on a 600 x 600 lattice: 246,488 particles, elastic
        return config_rings(600, precision)
    if name == "pse_500k":            # 1250 x 400 at 1.2 um, PSE, 120 m/min, jittered, 10000 steps
        return cut_config(1250, 400, 1.2, thermal="pse", speed_m_min=120.0, jitter=0.01, precision=precision)
    # calibration cuts: at 70 m/min the tool advances 0.04 dx in 400 steps and no particle
    # yields, so the force would not depend on the Johnson-Cook constants; at 1000 m/min the
    # contact zone yields within the first 50 steps
    if name == "adjoint_6k":
        return cut_config(150, 40, 10.0, thermal="adiabatic", speed_m_min=1000.0, precision=precision)
    if name == "adjoint_40k":
        return cut_config(375, 100, 4.0, thermal="adiabatic", speed_m_min=1000.0, precision=precision)
    raise KeyError(name)


@dataclass
class SimInfo:
    n_particles: int
    n_tool: int
    grid_nx: int
    grid_ny: int
    num_cell: int
    precision: int
    dt: float
    t_final: float
    mass: float
    bbmin_x: float
    bbmin_y: float
    cell_size: float
    step: int
    out_of_grid: int  # particles binned outside the grid so far (0 in a valid run)
    grid_overflow: int  # adaptive grid: steps whose geometry did not fit the cell table
    grid_adaptive: int
    cell_table: int   # cells of the binning table


class Sim:
    """One simulation handle: setup_ref_cut + leap_frog (+ report_force)."""

    def __init__(self, cfg: Config):
        self._h = C.c_void_p()
        _check(lib().mfc_create(C.byref(cfg), C.byref(self._h)), "mfc_create")
        self.cfg = cfg.copy()

    def close(self) -> None:
        if self._h:
            lib().mfc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def save(self, path: str) -> None:
        """Binary checkpoint of the handle (mfc_save)."""
        _check(lib().mfc_save(self._h, os.fsencode(path)), "mfc_save")

    @classmethod
    def load(cls, path: str) -> "Sim":
        """A new handle continuing a checkpoint (mfc_load)."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        _check(lib().mfc_load(os.fsencode(path), C.byref(self._h)), "mfc_load")
        self.cfg = Config()
        _check(lib().mfc_get_config(self._h, C.byref(self.cfg)), "mfc_get_config")
        return self

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def info(self) -> SimInfo:
        i = Info()
        _check(lib().mfc_get_info(self._h, C.byref(i)), "mfc_get_info")
        return SimInfo(*[getattr(i, n) for n, _ in Info._fields_])

    @property
    def N(self) -> int:
        return self.info.n_particles

    def step(self, n: int = 1) -> None:
        _check(lib().mfc_step(self._h, int(n)), "mfc_step")

    def sync(self) -> None:
        _check(lib().mfc_sync(self._h), "mfc_sync")

    PHASES = ("binning", "predict", "interact", "update", "force")  # "update": split interaction only

    def profile(self, nsteps: int) -> dict:
        """Mean ms per step of each phase, timed with HIP events on the handle's stream."""
        ms = np.zeros(len(self.PHASES))
        _check(lib().mfc_profile(self._h, int(nsteps), ms.ctypes.data_as(C.POINTER(C.c_double))), "mfc_profile")
        return dict(zip(self.PHASES, ms.tolist()))

    def tool_force(self) -> np.ndarray:
        F = np.zeros(2)
        _check(lib().mfc_tool_force(self._h, F.ctypes.data_as(C.POINTER(C.c_double))), "mfc_tool_force")
        return F

    def set_tuning(self, **kw) -> None:
        """Launch configuration (results are independent of it): split_interact=0|1,
        compact_pairs=0|1, stage_limit=0..STAGE_CAP (0: the full LDS stage),
        xcd_chunk=0..XCD_CHUNK_MAX, sort_fast=0..SORT_FAST_MAX, sort_list>=0 (0: no bound)."""
        t = Tuning()
        _check(lib().mfc_get_tuning(self._h, C.byref(t)), "mfc_get_tuning")
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(lib().mfc_set_tuning(self._h, C.byref(t)), "mfc_set_tuning")

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().mfc_get_tuning(self._h, C.byref(t)), "mfc_get_tuning")
        return {"split_interact": t.split_interact, "compact_pairs": t.compact_pairs, "stage_limit": t.stage_limit,
                "xcd_chunk": t.xcd_chunk, "sort_fast": t.sort_fast, "sort_list": t.sort_list}

    def vjp_forward(self, nsteps: int, checkpoint_every: int, w0: int, w1: int) -> np.ndarray:
        """Run nsteps steps keeping a checkpoint every checkpoint_every steps; returns the mean
        tool force over the steps [w0, w1) of the run."""
        F = np.zeros(2)
        _check(lib().mfc_vjp_forward(self._h, int(nsteps), int(checkpoint_every), int(w0), int(w1),
                                     F.ctypes.data_as(C.POINTER(C.c_double))), "mfc_vjp_forward")
        return F

    def vjp_backward(self, cot) -> np.ndarray:
        """cot . d(mean force)/d(theta) for theta = THETA, of the last vjp_forward."""
        c = np.ascontiguousarray(cot, dtype=np.float64)
        g = np.zeros(6)
        _check(lib().mfc_vjp_backward(self._h, c.ctypes.data_as(C.POINTER(C.c_double)),
                                      g.ctypes.data_as(C.POINTER(C.c_double))), "mfc_vjp_backward")
        return g

    def force_history(self, cap: int = 1 << 20) -> np.ndarray:
        n = C.c_int64()
        _check(lib().mfc_force_history(self._h, None, 0, C.byref(n)), "mfc_force_history")
        buf = np.zeros((max(1, min(cap, n.value)), 2))
        _check(lib().mfc_force_history(self._h, buf.ctypes.data_as(C.POINTER(C.c_double)), buf.shape[0], C.byref(n)),
               "mfc_force_history")
        return buf[: n.value].copy()

    def reset_force_history(self) -> None:
        _check(lib().mfc_force_history_reset(self._h), "mfc_force_history_reset")

    def field(self, name: str) -> np.ndarray:
        fid, nc = FIELDS[name]
        N = self.N
        dt = np.int32 if name in INT_FIELDS else np.float64
        out = np.zeros((N, nc) if nc > 1 else (N,), dtype=dt)
        _check(lib().mfc_get_field(self._h, fid, out.ctypes.data_as(C.c_void_p), out.nbytes), f"mfc_get_field({name})")
        return out

    def set_field(self, name: str, values: np.ndarray) -> None:
        fid, nc = FIELDS[name]
        dt = np.int32 if name in INT_FIELDS else np.float64
        v = np.ascontiguousarray(values, dtype=dt)
        _check(lib().mfc_set_field(self._h, fid, v.ctypes.data_as(C.c_void_p), v.nbytes), f"mfc_set_field({name})")

    def cells(self) -> np.ndarray:
        nc = self.info.num_cell
        out = np.zeros(nc + 1, dtype=np.int32)
        _check(lib().mfc_get_cells(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), nc + 1), "mfc_get_cells")
        return out

    def tool(self) -> dict:
        t = ToolState()
        _check(lib().mfc_get_tool(self._h, C.byref(t)), "mfc_get_tool")
        return t.as_arrays()

    def write_vtk(self, path: str) -> None:
        """The particles as the reference's legacy VTK file (vtk_writer_write)."""
        _check(lib().mfc_write_vtk(self._h, os.fsencode(path)), "mfc_write_vtk")

    def write_tool_vtk(self, path: str) -> None:
        """The tool as the reference's legacy VTK file of triangles."""
        _check(lib().mfc_write_tool_vtk(self._h, os.fsencode(path)), "mfc_write_tool_vtk")

    def tool_wear(self, dt: float = 0.0, K: float = 0.078, alpha: float = 2500.0, tool_vel=None) -> np.ndarray:
        """tool_wear::eval_usui over the current state: (min, max, mean) of the Usui wear rate
        over the particles in contact and the wear accumulated by this handle (sum of dt * mean).
        tool_vel defaults to (0, tool speed), as the reference's setup passes it."""
        if tool_vel is None:
            v = np.asarray(self.tool()["vel"], dtype=np.float64)
            tool_vel = (0.0, float(np.hypot(v[0], v[1])))
        tv = np.ascontiguousarray(tool_vel, dtype=np.float64)
        out = np.zeros(4)
        _check(lib().mfc_tool_wear(self._h, float(K), float(alpha), tv.ctypes.data_as(C.POINTER(C.c_double)),
                                   float(dt), out.ctypes.data_as(C.POINTER(C.c_double))), "mfc_tool_wear")
        return out


def device_math(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    yp = None
    if y is not None:
        y = np.ascontiguousarray(y, dtype=np.float64)
        yp = y.ctypes.data_as(C.POINTER(C.c_double))
    _check(lib().mfc_device_math(fn, x.ctypes.data_as(C.POINTER(C.c_double)), yp,
                                 out.ctypes.data_as(C.POINTER(C.c_double)), x.size), "mfc_device_math")
    return out


def scenario(cfg: Config) -> dict:
    """Host-side setup of a configuration (no device work): initial particles and tool."""
    info = Info()
    _check(lib().mfc_scenario(C.byref(cfg), C.byref(info), None, None, None, None, None, None), "mfc_scenario")
    N = info.n_particles
    pos = np.zeros((N, 2)); rho = np.zeros(N); h = np.zeros(N); T = np.zeros(N)
    flags = np.zeros(N, dtype=np.int32)
    t = ToolState()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _check(lib().mfc_scenario(C.byref(cfg), C.byref(info), dp(pos), dp(rho), dp(h), dp(T),
                              flags.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(t)), "mfc_scenario")
    return {"info": SimInfo(*[getattr(info, n) for n, _ in Info._fields_]), "pos": pos, "rho": rho, "h": h,
            "T": T, "flags": flags, "tool": t.as_arrays()}


class _Instance(Sim):
    """Read-out view of one instance of a Batch (owned by the batch)."""

    def __init__(self, handle, cfg, owner):
        self._h = handle
        self.cfg = cfg
        self._owner = owner  # keeps the batch alive

    def close(self) -> None:
        self._h = C.c_void_p()


class Batch:
    """n independent instances stepped together: one launch per phase for all of them."""

    def __init__(self, cfgs):
        arr = (Config * len(cfgs))(*cfgs)
        self._h = C.c_void_p()
        _check(lib().mfc_batch_create(arr, len(cfgs), C.byref(self._h)), "mfc_batch_create")
        self.cfgs = [c.copy() for c in cfgs]
        self._tuning = {"split_interact": 0, "compact_pairs": 0}
        self.instances = []
        for i, c in enumerate(self.cfgs):
            h = C.c_void_p()
            _check(lib().mfc_batch_instance(self._h, i, C.byref(h)), "mfc_batch_instance")
            self.instances.append(_Instance(h, c, self))

    def __len__(self):
        return len(self.instances)

    def __getitem__(self, i) -> Sim:
        return self.instances[i]

    def close(self) -> None:
        if self._h:
            for inst in self.instances:
                inst.close()
            lib().mfc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, **kw) -> None:
        """Launch configuration of every instance: split_interact=0|1, compact_pairs=0|1,
        stage_limit=0..STAGE_CAP."""
        self._tuning.update({k: int(v) for k, v in kw.items()})
        t = Tuning()
        for k, v in self._tuning.items():
            setattr(t, k, v)
        _check(lib().mfc_batch_set_tuning(self._h, C.byref(t)), "mfc_batch_set_tuning")

    def step(self, n: int = 1) -> None:
        _check(lib().mfc_batch_step(self._h, int(n)), "mfc_batch_step")

    def sync(self) -> None:
        _check(lib().mfc_batch_sync(self._h), "mfc_batch_sync")

    def vjp_forward(self, nsteps: int, checkpoint_every: int, w0: int, w1: int) -> np.ndarray:
        F = np.zeros((len(self), 2))
        _check(lib().mfc_batch_vjp_forward(self._h, int(nsteps), int(checkpoint_every), int(w0), int(w1),
                                           F.ctypes.data_as(C.POINTER(C.c_double))), "mfc_batch_vjp_forward")
        return F

    def vjp_backward(self, cot) -> np.ndarray:
        c = np.ascontiguousarray(cot, dtype=np.float64).reshape(len(self), 2)
        g = np.zeros((len(self), 6))
        _check(lib().mfc_batch_vjp_backward(self._h, c.ctypes.data_as(C.POINTER(C.c_double)),
                                            g.ctypes.data_as(C.POINTER(C.c_double))), "mfc_batch_vjp_backward")
        return g


def batch_configs(n: int = 32, seed: int = 0x355, precision: int = PREC_FP64) -> list:
    """Independent cuts drawn from a seed: depth of cut U[20, 100] um, speed U[30, 200] m/min,
    dx in {4, 6, 10} um, Johnson-Cook constants each scaled by U[0.9, 1.1], workpiece
    U[0.8, 2.0] mm x U[0.25, 0.5] mm (about 2k to 60k particles)."""
    rng = np.random.default_rng(seed)
    out = []
    A, B, nn, Cc, m = 1098.0, 1092.0, 0.93, 0.014, 1.1
    for _ in range(n):
        dx = float(rng.choice([4.0, 6.0, 10.0]))
        nx = int(round(rng.uniform(0.8, 2.0) * 1000.0 / dx))
        ny = int(round(rng.uniform(0.25, 0.5) * 1000.0 / dx))
        f = rng.uniform(0.9, 1.1, 5)
        jc = (A * f[0], B * f[1], nn * f[2], Cc * f[3], m * f[4])
        out.append(cut_config(nx, ny, dx, thermal="adiabatic", speed_m_min=float(rng.uniform(30.0, 200.0)),
                              feed_um=float(rng.uniform(20.0, 100.0)), jc=jc, precision=precision))
    return out

