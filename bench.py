"""Benchmark: ns/day of the MD core on one MI355X, with its roofline and the
CPU oracle timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 26,749-atom system (a 2,600-atom AMBER-typed chain, a 24-atom
ligand, 8 ions and 8,039 TIP3P waters in a truncated octahedron at 1.0 g/cm^3,
PME 1.0 nm cutoff, order 5, tolerance 5e-4, H-bond constraints + SETTLE,
LFMiddle Langevin 300 K / 1 ps^-1, dt 2 fs, fp32 arithmetic with Q32.32
int64 accumulation), built synthetically with a fixed seed, minimised, given
Maxwell-Boltzmann velocities and warmed up before the timed region.  The
plain single-system path runs strides of 50 steps captured as HIP graphs.

Prints one JSON line (rank 0).  Progress goes to stderr.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("making-it-rain-webgpu_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
PAIR_FLOPS = 76            # algorithmic FP32 flops per in-cutoff pair


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_synchronize():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def cpu_baseline(built, positions, velocities, budget_s):
    """The oracle (f64 C restatement of the reference's CPU path), one thread,
    LFMiddle steps on the same system and state until ~budget_s seconds."""
    import oracle as O
    o = O.OracleSystem(built.system)
    grid = o.default_grid(built.system.boxVectors)
    pos = positions.copy()
    vel = velocities.copy()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.run_lfmiddle(pos, vel, built.system.boxVectors, 1, 0.002, 300.0, 1.0, 1e-6, 1, steps, grid)
        steps += 1
        if time.perf_counter() - t0 >= budget_s or steps >= 50:
            break
    el = time.perf_counter() - t0
    ns_day = steps * 0.002e-3 / el * 86400.0
    return {"value": ns_day, "unit": "ns/day", "cores": 1, "kind": "port",
            "sample": f"{steps} LFMiddle steps (forces, SETTLE/SHAKE, thermostat) of the same "
                      f"{built.system.particleCount}-atom system by the f64 C oracle on one host core, "
                      f"{el:.1f} s", "seconds": el, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--minimise-iterations", type=int, default=300)
    ap.add_argument("--max-batch", type=int, default=0,
                    help="steps per captured HIP graph (0: 50; 1: launch every step directly)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build is single-GPU (asked for --gpus {args.gpus}). "
              "Independent replicas would shard across GPUs with no communication between them.",
              file=sys.stderr)
        sys.exit(2)

    from mdcore import synthetic
    from mdcore.engine import HipEngine, RunOptions

    t = time.perf_counter()
    built = synthetic.config(1)
    n = built.system.particleCount
    log(f"built {n} atoms ({built.description['shape']}, {built.description['width_nm']:.3f} nm) "
        f"in {time.perf_counter() - t:.1f} s")
    engine = HipEngine(built.system, built.state, RunOptions(timeStep=0.002, temperature=300.0, friction=1.0,
                                                            constraintTolerance=1e-6, seed=2026,
                                                            maxBatch=args.max_batch))
    t = time.perf_counter()
    from mdcore.engine import minimise
    res = minimise(engine, args.minimise_iterations, 10.0)
    log(f"minimised to {res.energy:.1f} kJ/mol in {res.iterations} iterations, {time.perf_counter() - t:.1f} s")
    engine.setVelocitiesToTemperature(300.0, 2026)
    t = time.perf_counter()
    engine.run(args.warmup)
    device_synchronize()
    rep = engine.report()
    log(f"warmed up {args.warmup} steps in {time.perf_counter() - t:.1f} s: T = {rep.temperature:.1f} K, "
        f"E = {rep.totalEnergy:.1f}")

    # ---- timed region ---------------------------------------------------
    device_synchronize()
    t0 = time.perf_counter()
    engine.run(args.steps)
    device_synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1000.0
    ns_day = args.steps * 0.002e-3 / elapsed * 86400.0
    rep = engine.report()
    log(f"timed {args.steps} steps: {ms_per_step * 1000:.1f} us/step, {ns_day:.1f} ns/day, "
        f"T = {rep.temperature:.1f} K")

    # ---- kernel roofline: per-phase HIP-event timing on the engine's stream --
    native = engine._native
    pairs, entries = native.pair_census()
    phases = native.profile_steps(args.profile_steps)
    pairs2, _ = native.pair_census()
    pairs_avg = 0.5 * (pairs + pairs2)
    t_pairs = phases["pairs"] * 1e-3
    achieved = pairs_avg * PAIR_FLOPS / t_pairs / 1e12 if t_pairs > 0 else 0.0
    step_ms = sum(phases.values())
    log("phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()))
    # the fused integrator is HBM-bound: 124 algorithmic bytes per atom-step
    integ_bytes = 124.0 * n
    integ_gbs = integ_bytes / (phases["integrate"] * 1e-3) / 1e9 if phases["integrate"] > 0 else 0.0
    flags = engine.verletFlags()

    result = {
        "metric": "aggregate ns/day at 2 fs (batched+learned vs plain, same GPU, fixed force error)",
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 arithmetic, int64 Q32.32 accumulation",
        "data": "synthetic (seeded generator: chain + ligand + TIP3P + ions), minimised and thermalised",
        "config": {
            "workload": "plain single-system path: 26,749 atoms, truncated octahedron "
                        f"{built.description['width_nm']:.3f} nm, PME 1.0 nm cutoff, LFMiddle 300 K, dt 2 fs, "
                        "H-bond SHAKE + SETTLE, 50-step HIP-graph strides",
            "atoms": n,
            "atom_steps_per_s": n * args.steps / elapsed,
            "in_cutoff_pairs": int(pairs_avg),
            "list_entries": int(entries),
            "list_rebuilds": flags["rebuilds"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "k_nb_pairs (LJ + erfc-Ewald pair kernel)",
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic": f"{PAIR_FLOPS} flop x {int(pairs_avg)} in-cutoff pairs per launch",
            "kernel_ms": phases["pairs"],
        },
        "phases_ms": phases,
        "profiled_step_ms": step_ms,
        "integrator_roofline": {"bound": "hbm", "achieved": integ_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": integ_gbs / HBM_PEAK_GBS, "bytes_per_atom": 124},
    }
    if not args.no_cpu_baseline:
        log(f"timing the CPU oracle for ~{args.cpu_seconds:.0f} s")
        result["cpu_baseline"] = cpu_baseline(built, engine.positions(), engine.velocities(), args.cpu_seconds)
        log(f"cpu baseline {result['cpu_baseline']['value']:.4f} ns/day")
    engine.destroy()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
