"""Writes the small golden fixtures of tests/golden/ from the FP64 oracle.

The reference ships no tests or vectors; these fixtures freeze the oracle's outputs (itself pinned
by the closed-form, quadrature and brute-force tests) so that later changes to the oracle or to
the HIP path are caught bit for bit:
  binning_12x8_jitter.npz   cell ids, sorted order and cell offsets of a jittered 12 x 8 lattice
  trace_20x10_tool.npz      5-step state trace of a 20 x 10 cut with tool particles and a fillet
  jc_return_table.npz       single-particle Johnson-Cook radial returns
  contact_cases.npz         penalty contact and friction of points around the tool
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "mfree_iwf-ul_cut_gpu_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import pymfc  # noqa: E402
import oracle_binding as ob  # noqa: E402

JC = np.array([0.01098, 0.01092, 0.014, 0.93, 1.1, 1836.0, 300.0, 1e-6])
TRACE_FIELDS = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "id", "flags")


def cfg_binning():
    return pymfc.cut_config(12, 8, 10.0, jitter=0.45, seed=12)


def cfg_trace():
    c = pymfc.cut_config(20, 10, 10.0, thermal="contact")
    return c


def main():
    o = ob.OracleSim(cfg_binning())
    rec = {"pos0": o.field("pos")}
    for s in (1, 2):
        o.step(1)
        rec[f"hash{s}"], rec[f"id{s}"], rec[f"cells{s}"] = o.field("hash"), o.field("id"), o.cells()
    np.savez_compressed(os.path.join(HERE, "binning_12x8_jitter.npz"), **rec)

    o = ob.OracleSim(cfg_trace())
    rec = {"n_particles": o.N}
    for s in range(1, 6):
        o.step(1)
        for f in TRACE_FIELDS:
            rec[f"{f}_{s}"] = o.field(f)
    rec["force"] = o.force_history()
    np.savez_compressed(os.path.join(HERE, "trace_20x10_tool.npz"), **rec)

    rng = np.random.default_rng(2024)
    ins, outs = [], []
    for _ in range(64):
        S = rng.normal(0, 0.012, 4)
        st = np.array([*S, rng.normal(0, 0.003), rng.uniform(0, 0.8), 10 ** rng.uniform(-7, -2),
                       rng.uniform(300, 1200), 4.43 * (1 + rng.normal(0, 1e-3))])
        ins.append(st)
        outs.append(ob.radial_return(JC, 0.424, 8e-4, 0.9, 553e-8, st))
    np.savez_compressed(os.path.join(HERE, "jc_return_table.npz"), jc=JC, G=0.424, dt=8e-4, tq=0.9, cp=553e-8,
                        inputs=np.array(ins), outputs=np.array(outs))

    o = ob.OracleSim(pymfc.workload("adiabatic_6k"))
    t = o.tool()
    cx, cy, r = t["fillet"][:3]
    fx = t["left"][1][0]
    pts = np.array([[fx - 2e-5, 0.07], [fx + 2e-5, 0.07], [cx + 0.3 * r, cy - 0.8 * r], [cx - 0.9 * r, cy - 0.2 * r],
                    [fx - 1e-4, 0.0501], [cx, cy - 0.999 * r], [fx - 3e-4, 0.0505], [-0.02, 0.0502]])
    vel = np.array([[1.2e-4, 0.0], [0.0, 1e-5], [1e-4, 3e-6], [0.0, 0.0], [2e-4, -1e-4], [1e-4, 1e-4], [0.0, 0.5],
                    [1e-4, 0.0]])
    fold = np.array([[0.0, 0.0], [0.0, 0.0], [1e-10, -1e-10], [0.0, 0.0], [0.0, 1e-9], [0.0, 0.0], [0.0, 0.0],
                     [1e-9, 1e-9]])
    out = np.array([o.contact_point(p[0], p[1], v[0], v[1], f[0], f[1]) for p, v, f in zip(pts, vel, fold)])
    np.savez_compressed(os.path.join(HERE, "contact_cases.npz"), points=pts, vel=vel, ft_old=fold, out=out)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
