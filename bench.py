"""Benchmark of the MI355X SPH cutting step: particle-steps/s on one GPU.

Workload (default): the thermal-contact orthogonal cut at ~150k particles — a 750 x 200
Ti6Al4V lattice at dx = 2 um with Brookshaw conduction and 46,979 sampled tool particles
(196,979 particles in total). Synthetic: the lattice and tool are generated, nothing is read.

One step = one leap_frog::step of the reference (leap_frog.cu:94-155) over all particles.
value = particles x timed steps / wall seconds, with the state resident in HBM.

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
FP64_VALU_PEAK_TFLOPS = 78.6

# compulsory bytes per particle of one launch of the interaction kernel (FP64): the 128-byte
# neighbour record of the particle (neighbour re-reads are cache traffic), 26 state doubles
# read and written, flags read and written, the positional p (read) and contact normal (read,
# written), the sorted cell id (read)
INTERACT_BYTES_FP64 = 128 + 26 * 8 + 4 + 8 + 16 + 4 + (26 * 8 + 4 + 16)
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="thermal_contact_150k")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not an implementation)", file=sys.stderr)
        return 2

    import pymfc

    cfg = pymfc.workload(args.workload, pymfc.PREC_FP64)
    sim = pymfc.Sim(cfg)
    info = sim.info
    N = info.n_particles
    print(f"[bench] {args.workload}: {N} particles ({info.n_tool} tool), {info.num_cell} cells, dt {info.dt:.4e} us",
          file=sys.stderr, flush=True)
    sim.step(args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    value = N * args.steps / wall
    print(f"[bench] fp64: {args.steps} steps in {wall:.3f} s -> {value:.4e} particle-steps/s", file=sys.stderr,
          flush=True)
    ph = sim.profile(args.profile_steps)
    print(f"[bench] phases (ms/step): {ph}", file=sys.stderr, flush=True)
    t_int = ph["interact"] * 1e-3
    achieved = N * INTERACT_BYTES_FP64 / t_int / 1e9
    F = sim.tool_force()
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_particles": N, "n_tool": info.n_tool, "num_cell": info.num_cell,
                   "precision": "fp64", "thermal": "brookshaw+tool-contact"},
        "roofline": {"bound": "hbm", "kernel": "k_interact", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_particle": INTERACT_BYTES_FP64, "ms_per_launch": ph["interact"]},
        "phases_ms": ph,
        "step_hbm_gbs": N * STEP_BYTES_FP64 / (wall / args.steps) / 1e9,
        "tool_force": F.tolist(),
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
