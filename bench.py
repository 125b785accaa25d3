"""Benchmark of the MI355X SPH cutting step: particle-steps/s on one GPU, FP32-compensated
path against the plain FP64 path.

Workload (default): the thermal-contact orthogonal cut at ~150k particles — a 750 x 200
Ti6Al4V lattice at dx = 2 um with Brookshaw conduction and 46,979 sampled tool particles
(196,979 particles in total). Synthetic: the lattice and tool are generated, nothing is read.

One step = one leap_frog::step of the reference (leap_frog.cu:94-155) over all particles.
value = particles x timed steps / wall seconds of the FP32-compensated path, state resident
in HBM; the FP64 path is timed the same way and reported beside it. By default the timed steps
are the cut's 20,000 steps (after 116 warm-up steps): the cost of a step grows as the chip forms
and more particles yield, so a short window early in the cut would overstate the rate.

Printed: one JSON line with the metric, the dominant kernel's roofline (HIP-event timed),
and the FP64 host oracle timed on this machine's cores on a bounded sample (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mfree_iwf-ul_cut_gpu_amd"))

METRIC = "particle-steps/s at 150k particles (FP32-comp vs FP64); adjoint/forward ratio"
THERMAL_OF_WORKLOAD = {"thermal_contact_150k": "brookshaw+tool-contact", "pse_500k": "pse", "adiabatic_6k": "adiabatic",
                       "rings_250k": "adiabatic (elastic rings)"}
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle-step of a fully fused step, counted in the
# reference's storage (FP64 fields, flags as three doubles): the read-only set (h 8, flags 24)
# and the mutable set (pos 16, vel 16, S 32, rho 8, T 8, eps_pl 8, eps_pl_dot 8, pos_t 16,
# vel_t 16, S_t 32, rho_t 8, T_t 8, ft 16, fc 16, n 16 = 224) read once, the mutable set written
# once, and 12 B of binning keys (cell id written and read, index): 256 + 224 + 12 = 492 B.
STEP_BYTES = 256 + 224 + 12
# The same 492 B split over the step's kernels, each compulsory byte counted in the first kernel
# that touches it: the binning passes take the 12 B of keys; predict reads the read-only set and
# the fields of the leap-frog predictor (h, flags, pos, vel, S, rho, T and their rates: 192 B);
# the interaction kernel reads the rest of the mutable set (eps_pl, eps_pl_dot, fc, ft, n: 64 B)
# and writes the whole mutable set (224 B). Everything else a kernel moves (neighbour records,
# the start state read again by the update, cell tables) is traffic above the algorithmic bytes.
KERNEL_BYTES = {"binning": 12, "predict": 192, "interact": 64 + 224}
assert sum(KERNEL_BYTES.values()) == STEP_BYTES
# the split variant: the sums kernel reads nothing compulsory beyond what the interaction kernel
# is credited with; its 288 B are credited to the pair (sums + update)
# HBM bytes per particle of one launch of the fused interaction kernel on thermal_contact_150k,
# by precision, pair loop, XCD chunk length and LDS stage limit (0: the full 2048 candidates),
# measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes and scaled by a 512 MiB copy (the
# commands and tables are in profiles/traffic_thermal_contact_150k.md); a table of measurements,
# not measured in this run, and reported as `traffic` only for the tuning it was measured with
MEASURED_TRAFFIC = {("kb_interact", "fp32c", 0, 8, 0): 664, ("kb_interact", "fp32c", 0, 32, 0): 612,
                    ("kb_interact", "fp32c", 1, 8, 0): 647, ("kb_interact", "fp32c", 1, 32, 0): 599,
                    ("kb_interact", "fp64", 0, 8, 0): 826, ("kb_interact", "fp64", 0, 32, 0): 743,
                    ("kb_interact", "fp64", 1, 8, 0): 794, ("kb_interact", "fp64", 1, 32, 0): 691,
                    ("kb_interact", "fp32c", 0, 8, 1536): 669, ("kb_interact", "fp32c", 0, 32, 1536): 615,
                    ("kb_interact", "fp32c", 1, 8, 1536): 656, ("kb_interact", "fp32c", 1, 32, 1536): 600,
                    ("kb_interact", "fp64", 0, 8, 1536): 846, ("kb_interact", "fp64", 0, 32, 1536): 748,
                    ("kb_interact", "fp64", 1, 8, 1536): 826, ("kb_interact", "fp64", 1, 32, 1536): 695}
MEASURED_TRAFFIC_SOURCE = ("table: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, thermal_contact_150k, fused interaction, "
                           "profiles/traffic_thermal_contact_150k.md")


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


FIXED_TUNING = {}  # --tuning k=v,...: the settings used when the autotuner is off


def tune(pymfc, handle, steps: int) -> dict:
    """The package's autotuner on a live handle (every launch configuration gives the same bits);
    off when steps is 0: the defaults, or the settings given with --tuning."""
    if steps <= 0:
        return {k: v[0] for k, v in pymfc.TUNING_SPACE.items()} | FIXED_TUNING | {"ms_per_step": {}}
    return pymfc.autotune(handle, steps)


def run_path(pymfc, workload: str, precision: int, steps: int, warmup: int, profile_steps: int,
             tune_steps: int) -> dict:
    """The timed steps are steps warmup + 16 .. warmup + 16 + steps of the run whatever the tuning:
    the autotuner runs on a handle of its own, advanced to the middle of the timed window (the
    cost of a step, and which launch settings are fastest, change along the cut as the chip forms
    and more particles yield: the tuner's candidates then run through the window's second half),
    and the measured handle takes its choice, runs warmup steps and 2 * GRAPH_STEPS more (its
    hipGraphs are captured there), then the timed steps."""
    cfg = pymfc.workload(workload, precision)
    tuning = {"ms_per_step": {}}
    if tune_steps > 0:
        t = pymfc.Sim(cfg)
        t.step(warmup + 2 * pymfc.GRAPH_STEPS + steps // 2)
        tuning = tune(pymfc, t, tune_steps)
        t.close()
    else:
        tuning = tune(pymfc, None, 0)
    sim = pymfc.Sim(cfg)
    sim.set_tuning(**{k: tuning[k] for k in pymfc.TUNING_SPACE})
    info = sim.info
    N = info.n_particles
    sim.step(warmup)
    sim.step(2 * pymfc.GRAPH_STEPS)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ph = sim.profile(profile_steps) if profile_steps > 0 else None
    F = sim.tool_force()
    sim.close()
    return {"N": N, "info": info, "wall": wall, "value": N * steps / wall, "phases": ph, "force": F.tolist(),
            "tuning": tuning}


def other_configs(pymfc, steps: int, tune_steps: int) -> dict:
    """The other single-GPU configurations: the 6k adiabatic cut (150 x 40 at 10 um), the 500k
    PSE cut (1250 x 400 at 1.2 um, jittered) and the colliding rings of setup_rings on a 600 x 600
    lattice (246k particles, no tool, elastic), both precisions, each autotuned."""
    out = {}
    for wl in ("adiabatic_6k", "pse_500k", "rings_250k"):
        out[wl] = {}
        for name, prec in (("fp64", pymfc.PREC_FP64), ("fp32c", pymfc.PREC_FP32C)):
            r = run_path(pymfc, wl, prec, steps, 20, 10, tune_steps)
            out[wl][name] = {"n_particles": r["N"], "particle_steps_per_s": r["value"],
                             "ms_per_step": r["wall"] / steps * 1e3, "tuning": {k: v for k, v in r["tuning"].items()
                                                                               if k != "ms_per_step"}}
        out[wl]["fp32c_over_fp64"] = out[wl]["fp32c"]["particle_steps_per_s"] / out[wl]["fp64"]["particle_steps_per_s"]
    return out


def adjoint_ratio(pymfc, workload: str, nsteps: int, K: int) -> dict:
    """Reverse mode on the calibration workload: wall time of the backward pass (replay of the
    checkpoint segments included) over the plain forward of the same steps."""
    cfg = pymfc.workload(workload, pymfc.PREC_FP64)
    cfg.promote_fp64 = 1
    cfg.secant_tol = 1e-14
    f = pymfc.Sim(cfg)
    f.step(20)
    f.sync()
    t0 = time.perf_counter()
    f.step(nsteps)
    f.sync()
    t_fwd = time.perf_counter() - t0
    N = f.N
    f.close()
    a = pymfc.Sim(cfg)
    a.step(20)
    a.sync()
    t0 = time.perf_counter()
    a.vjp_forward(nsteps, K, 0, nsteps)
    t_tape = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = a.vjp_backward([1.0, 0.0])
    t_bwd = time.perf_counter() - t0
    a.close()
    return {"workload": workload, "n_particles": N, "steps": nsteps, "checkpoint_every": K,
            "forward_ms_per_step": t_fwd / nsteps * 1e3, "taped_forward_ms_per_step": t_tape / nsteps * 1e3,
            "backward_ms_per_step": t_bwd / nsteps * 1e3, "ratio": t_bwd / t_fwd,
            "ratio_with_tape": (t_tape + t_bwd) / t_fwd, "grad_Fx": g.tolist()}


def batch_bench(pymfc, n: int, steps: int, tune_steps: int) -> dict:
    """C5: n independent uneven cuts in one batch (one launch per phase for all instances),
    against the same instances run one after another as single handles, over the same steps of
    the run (10 + 2 * GRAPH_STEPS warm-up steps, the autotuner on a batch of its own advanced to
    the middle of the timed window)."""
    out = {"instances": n}
    warm = 10 + 2 * pymfc.GRAPH_STEPS
    for name, prec in (("fp64", pymfc.PREC_FP64), ("fp32c", pymfc.PREC_FP32C)):
        cfgs = pymfc.batch_configs(n, precision=prec)
        if tune_steps > 0:
            tb = pymfc.Batch(cfgs)
            tb.step(warm + steps // 2)  # the middle of the timed window, as run_path
            tuning = tune(pymfc, tb, max(tune_steps, 2))
            tb.close()
        else:
            tuning = tune(pymfc, None, 0)
        best = {k: tuning[k] for k in pymfc.TUNING_SPACE}
        b = pymfc.Batch(cfgs)
        b.set_tuning(**best)
        N = sum(inst.N for inst in b.instances)
        b.step(warm)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        wall = time.perf_counter() - t0
        sizes = [inst.N for inst in b.instances]
        b.close()
        seq = 0.0
        for c in cfgs:
            s = pymfc.Sim(c)
            s.set_tuning(**best)
            s.step(warm)
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            seq += time.perf_counter() - t0
            s.close()
        out[name] = {"particles": N, "min_instance": min(sizes), "max_instance": max(sizes),
                     "particle_steps_per_s": N * steps / wall, "ms_per_step": wall / steps * 1e3,
                     "sequential_particle_steps_per_s": N * steps / seq, "batch_speedup": seq / wall,
                     "tuning": tuning}
    out["adjoint"] = batch_adjoint_bench(pymfc, n)
    return out


def batch_adjoint_bench(pymfc, n: int, nsteps: int = 100, K: int = 25) -> dict:
    """C5's reverse half: the taped forward and the backward pass of the n instances as one batch
    (one launch per phase over all instances' tiles), against the same instances one after
    another; promoted FP64 with a tight secant, a window over all nsteps steps."""
    cfgs = []
    for c in pymfc.batch_configs(n):
        c.promote_fp64 = 1
        c.secant_tol = 1e-14
        cfgs.append(c)
    cot = np.tile([1.0, 0.0], (n, 1))
    b = pymfc.Batch(cfgs)
    N = sum(inst.N for inst in b.instances)
    b.step(4)
    b.sync()
    t0 = time.perf_counter()
    b.vjp_forward(nsteps, K, 0, nsteps)
    t_tape = time.perf_counter() - t0
    t0 = time.perf_counter()
    G = b.vjp_backward(cot)
    t_bwd = time.perf_counter() - t0
    b.close()
    seq_tape = seq_bwd = 0.0
    for c in cfgs:
        s = pymfc.Sim(c)
        s.step(4)
        s.sync()
        t0 = time.perf_counter()
        s.vjp_forward(nsteps, K, 0, nsteps)
        seq_tape += time.perf_counter() - t0
        t0 = time.perf_counter()
        s.vjp_backward(cot[0])
        seq_bwd += time.perf_counter() - t0
        s.close()
    return {"particles": N, "steps": nsteps, "checkpoint_every": K, "tape_ms_per_step": t_tape / nsteps * 1e3,
            "backward_ms_per_step": t_bwd / nsteps * 1e3, "sequential_tape_ms_per_step": seq_tape / nsteps * 1e3,
            "sequential_backward_ms_per_step": seq_bwd / nsteps * 1e3, "backward_speedup": seq_bwd / t_bwd,
            "grad_Fx_instance0": G[0].tolist()}


COPY_PROBE_BYTES = 512 << 20


def counter_probe(pymfc, workload: str, precision: int, steps: int, split: int, compact: int) -> None:
    """A short run for hardware-counter passes (rocprofv3 --pmc): `steps` steps of one path,
    then a plain device copy of COPY_PROBE_BYTES read + COPY_PROBE_BYTES written, whose counter
    readings calibrate the bytes per counted unit."""
    import torch
    sim = pymfc.Sim(pymfc.workload(workload, precision))
    sim.set_tuning(**(FIXED_TUNING | {"split_interact": split, "compact_pairs": compact}))
    for _ in range(steps):
        sim.step(1)
    sim.sync()
    a = torch.empty(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": workload, "precision": precision, "steps": steps, "split_interact": split,
                      "compact_pairs": compact,
                      "copy_bytes": COPY_PROBE_BYTES, "N": sim.N}), flush=True)
    sim.close()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="thermal_contact_150k")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32c"], default=None)
    ap.add_argument("--tune-steps", type=int, default=100, help="steps per candidate of the autotuner (0: off)")
    ap.add_argument("--tuning", default="", help="with --tune-steps 0: launch settings k=v,... (e.g. compact_pairs=1)")
    ap.add_argument("--adjoint-workload", default="adjoint_40k")
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-k", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=32)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--other-steps", type=int, default=500, help="timed steps of the other configurations")
    ap.add_argument("--no-other", action="store_true", help="skip the 6k and 500k configurations")
    ap.add_argument("--probe", action="store_true", help="counter-pass mode: short run + calibration copy")
    ap.add_argument("--split", type=int, default=0, help="split_interact for --probe")
    ap.add_argument("--compact", type=int, default=0, help="compact_pairs for --probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not an implementation)", file=sys.stderr)
        return 2

    import pymfc

    for kv in filter(None, args.tuning.split(",")):
        k, v = kv.split("=")
        if k not in pymfc.TUNING_SPACE:
            print(f"bench.py: unknown launch setting {k} (one of {list(pymfc.TUNING_SPACE)})", file=sys.stderr)
            return 2
        FIXED_TUNING[k] = int(v)
    if args.probe:
        counter_probe(pymfc, args.workload, pymfc.PREC_FP32C if args.only == "fp32c" else pymfc.PREC_FP64,
                      args.steps, args.split, args.compact)
        return 0

    res = {}
    for name, prec in (("fp64", pymfc.PREC_FP64), ("fp32c", pymfc.PREC_FP32C)):
        if args.only and args.only != name:
            continue
        r = run_path(pymfc, args.workload, prec, args.steps, args.warmup, args.profile_steps, args.tune_steps)
        res[name] = r
        print(f"[bench] {name}: {r['N']} particles ({r['info'].n_tool} tool), {args.steps} steps in {r['wall']:.3f} s "
              f"-> {r['value']:.4e} particle-steps/s; phases ms/step {r['phases']}; tuning {r['tuning']}",
              file=sys.stderr, flush=True)
    head = "fp32c" if "fp32c" in res else "fp64"
    r = res[head]
    N = r["N"]
    split = r["tuning"]["split_interact"]
    # the dominant kernel: the fused interaction, or the sums + update pair of the split variant
    kname = "kb_sums+kb_update" if split else "kb_interact"
    kbytes = KERNEL_BYTES["interact"]
    # --profile-steps 0 (kernel-trace runs) skips the event-timed phases, and with them the
    # kernel's roofline
    t_int = (r["phases"]["interact"] + r["phases"]["update"]) * 1e-3 if r["phases"] else None
    achieved = N * kbytes / t_int / 1e9 if t_int else None
    tkey = ("kb_interact", head, r["tuning"]["compact_pairs"], r["tuning"].get("xcd_chunk", 8),
            r["tuning"]["stage_limit"])
    traffic = None
    if args.workload == "thermal_contact_150k" and not split and tkey in MEASURED_TRAFFIC:
        traffic = MEASURED_TRAFFIC[tkey] * N
    step_s = r["wall"] / args.steps
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
                   "split_interact": split, "compact_pairs": r["tuning"]["compact_pairs"],
                   "num_cell": r["info"].num_cell, "precision": head,
                   "thermal": THERMAL_OF_WORKLOAD.get(args.workload, "see pymfc.workload")},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": MEASURED_TRAFFIC_SOURCE if traffic is not None else None,
                     "bytes_per_particle": kbytes, "ms_per_launch": t_int * 1e3 if t_int else None,
                     "step": {"bytes_per_particle": STEP_BYTES, "achieved": N * STEP_BYTES / step_s / 1e9,
                              "frac": N * STEP_BYTES / step_s / 1e9 / HBM_PEAK_GBS,
                              "kernel_bytes_per_particle": KERNEL_BYTES}},
        "tuning": r["tuning"],
        "step_hbm_gbs": N * STEP_BYTES / step_s / 1e9,
        "phases_ms": r["phases"],
        "tool_force": r["force"],
    }
    if "fp64" in res:
        result["fp64"] = {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["wall"] / args.steps * 1e3,
                          "phases_ms": res["fp64"]["phases"], "tuning": res["fp64"]["tuning"],
                          "step_hbm_gbs": N * STEP_BYTES / (res["fp64"]["wall"] / args.steps) / 1e9}
        if head == "fp32c":
            result["fp32c_over_fp64"] = res["fp32c"]["value"] / res["fp64"]["value"]
    result["adjoint_over_forward"] = None
    if not args.no_adjoint:
        adj = adjoint_ratio(pymfc, args.adjoint_workload, args.adjoint_steps, args.adjoint_k)
        print(f"[bench] adjoint: {adj}", file=sys.stderr, flush=True)
        result["adjoint_over_forward"] = adj["ratio"]
        result["adjoint"] = adj
    if not args.no_other and args.workload == "thermal_contact_150k":
        oc = other_configs(pymfc, args.other_steps, args.tune_steps)
        print(f"[bench] other configurations: {oc}", file=sys.stderr, flush=True)
        result["other_configs"] = oc
    if not args.no_batch:
        bb = batch_bench(pymfc, args.batch_instances, args.batch_steps, args.tune_steps)
        print(f"[bench] batch: {bb}", file=sys.stderr, flush=True)
        result["batch"] = bb
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(pymfc.workload(args.workload, pymfc.PREC_FP64), args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
