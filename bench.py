"""Benchmark of the MI355X SPH cutting step: particle-steps/s on one GPU, FP32-compensated
path against the plain FP64 path.

Workload (default): the thermal-contact orthogonal cut at ~150k particles — a 750 x 200
Ti6Al4V lattice at dx = 2 um with Brookshaw conduction and 46,979 sampled tool particles
(196,979 particles in total). Synthetic: the lattice and tool are generated, nothing is read.

One step = one leap_frog::step of the reference (leap_frog.cu:94-155) over all particles.
value = particles x timed steps / wall seconds of the FP32-compensated path, state resident
in HBM; the FP64 path is timed the same way and reported beside it.

Printed: one JSON line with the metric, the dominant kernel's roofline (HIP-event timed),
and the FP64 host oracle timed on this machine's cores on a bounded sample (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mfree_iwf-ul_cut_gpu_amd"))

METRIC = "particle-steps/s at 150k particles (FP32-comp vs FP64); adjoint/forward ratio"
HBM_PEAK_GBS = 8000.0

# compulsory bytes per particle of one launch of the interaction kernel: the particle's own
# neighbour record (neighbour re-reads are cache traffic), 26 state doubles read and written,
# flags read and written, the positional pressure (read) and contact normal (read, written),
# the sorted cell id (read)
_STATE_RW = 26 * 8 + 4 + 8 + 16 + 4 + (26 * 8 + 4 + 16)
INTERACT_BYTES = {"fp64": 128 + _STATE_RW, "fp32c": 64 + 16 + _STATE_RW}
# whole step, compulsory traffic of a fully fused FP64 step (persistent read-only 32 B,
# mutable 224 B read + written, 12 B of binning keys)
STEP_BYTES_FP64 = 256 + 224 + 12


def cpu_baseline(cfg, target_s: float) -> dict:
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_binding as ob
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    ob.set_threads(ncpu)
    o = ob.OracleSim(cfg)
    N = o.N
    o.step(1)  # warm-up, excluded
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = max(1, min(200, int(target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    wall = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": N * n / wall, "unit": "particle-steps/s", "cores": ob.threads(), "kind": "port",
            "sample": f"{n} steps of the same {N}-particle workload after 2 warm-up steps, FP64 OpenMP oracle, "
                      f"{wall:.1f} s wall on {ncpu} threads ({model})"}


def run_path(pymfc, workload: str, precision: int, steps: int, warmup: int, profile_steps: int) -> dict:
    cfg = pymfc.workload(workload, precision)
    sim = pymfc.Sim(cfg)
    info = sim.info
    N = info.n_particles
    sim.step(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ph = sim.profile(profile_steps)
    F = sim.tool_force()
    sim.close()
    return {"N": N, "info": info, "wall": wall, "value": N * steps / wall, "phases": ph, "force": F.tolist()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="thermal_contact_150k")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32c"], default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not an implementation)", file=sys.stderr)
        return 2

    import pymfc

    res = {}
    for name, prec in (("fp64", pymfc.PREC_FP64), ("fp32c", pymfc.PREC_FP32C)):
        if args.only and args.only != name:
            continue
        r = run_path(pymfc, args.workload, prec, args.steps, args.warmup, args.profile_steps)
        res[name] = r
        print(f"[bench] {name}: {r['N']} particles ({r['info'].n_tool} tool), {args.steps} steps in {r['wall']:.3f} s "
              f"-> {r['value']:.4e} particle-steps/s; phases ms/step {r['phases']}", file=sys.stderr, flush=True)
    head = "fp32c" if "fp32c" in res else "fp64"
    r = res[head]
    N = r["N"]
    t_int = r["phases"]["interact"] * 1e-3
    achieved = N * INTERACT_BYTES[head] / t_int / 1e9
    result = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated neighbour sums) + f64 state" if head == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_particles": N, "n_tool": r["info"].n_tool,
                   "num_cell": r["info"].num_cell, "precision": head, "thermal": "brookshaw+tool-contact"},
        "roofline": {"bound": "hbm", "kernel": "k_interact", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_particle": INTERACT_BYTES[head], "ms_per_launch": r["phases"]["interact"]},
        "phases_ms": r["phases"],
        "tool_force": r["force"],
    }
    if "fp64" in res:
        result["fp64"] = {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["wall"] / args.steps * 1e3,
                          "phases_ms": res["fp64"]["phases"],
                          "step_hbm_gbs": N * STEP_BYTES_FP64 / (res["fp64"]["wall"] / args.steps) / 1e9}
        if head == "fp32c":
            result["fp32c_over_fp64"] = res["fp32c"]["value"] / res["fp64"]["value"]
    result["adjoint_over_forward"] = None
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(pymfc.workload(args.workload, pymfc.PREC_FP64), args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
