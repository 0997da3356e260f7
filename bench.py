"""Benchmark: aggregate ns/day of the MD core on one MI355X, batched vs plain,
with the dominant kernel's roofline and the CPU oracle timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ensemble|single]

Workload "ensemble" (the default; the metric's configuration): 64 independent
solvated chains whose atom counts are log-uniform on 2,000-30,000 (647,236
atoms in all), cubic and truncated-octahedral cells alternating, per-replica
seeds and temperatures 290-320 K, PME 1.0 nm cutoff, order 5, tolerance 5e-4,
H-bond constraints + SETTLE, LFMiddle Langevin 1 ps^-1, dt 2 fs, fp32
arithmetic with Q32.32 int64 accumulation.  Built synthetically from a fixed
seed, minimised, given Maxwell-Boltzmann velocities, equilibrated for 400
steps of 0.5 fs at 10/ps friction (straight from the minimiser to 2 fs,
strained spots of a generated chain can blow up) and warmed up at 2 fs.
Then the same replicas are timed three ways for K steps each:

* plain: one after another, each through its own single-system engine
  (strides of 50 steps captured as HIP graphs);
* batched: all 64 together, one kernel launch per phase for every replica,
  strides captured as HIP graphs, plain SHAKE;
* batched + learned: the same, after a 32-64-64-16 network has been trained
  on SHAKE multipliers recorded from every replica (outside the timed
  region; its cost is reported) and attached to every replica, so each
  constraint cluster's SHAKE starts from the network's guess.

`value` is the batched + learned aggregate ns/day: (sum over replicas of K *
dt) / wall time.  Constraints converge to the same tolerance either way; the
forces of sampled replicas are checked against the f64 oracle after the run
(the metric's fixed force error, 2e-6 relative RMS) and every constrained
distance against its length.

Workload "single": the 26,749-atom system alone (truncated octahedron), the
plain path's single-system throughput.

Prints one JSON line.  Progress goes to stderr.
"""
import argparse
import json
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
# HBM bytes per launch of the batched pair kernel on this ensemble, by the
# work-unit length in use (the autotuner picks 64 or 128 entries here):
# FETCH_SIZE + WRITE_SIZE from two rocprofv3 --pmc passes over
# `bench.py --traffic-probe 3 --pair-segment N`, each scaled by its reading
# for a 256 MiB device copy (profiles/k_nb_pairs_ensemble64_traffic.md)
PAIR_TRAFFIC_BYTES = {64: 465.0e6, 128: 436.7e6}
DT_PS = 0.002
METRIC = "aggregate ns/day at 2 fs (batched+learned vs plain, same GPU, fixed force error)"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_synchronize():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def _build_one(plan):
    from mdcore import synthetic
    b = synthetic.build(plan["n_chain"], plan["n_waters"], plan["n_ions"], plan["shape"], plan["seed"],
                        grid=plan["grid"])
    return synthetic.Replica(b, plan["seed"], plan["temperature"])


def build_ensemble(count, workers):
    """The ensemble's replicas, built in spawned worker processes (no GPU is
    touched before they finish)."""
    from mdcore import synthetic
    plan = synthetic.ensemble_plan(count)
    if workers <= 1:
        return [_build_one(p) for p in plan]
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_build_one, plan, chunksize=1)


def oracle_rate(system, positions, velocities, budget_s, temperature=300.0, seed=1):
    """Atom-steps/s of the f64 C oracle (the reference CPU path restated) on
    one host core: LFMiddle steps with SETTLE/SHAKE and the thermostat."""
    import oracle as O
    o = O.OracleSystem(system)
    stated = tuple(int(g) for g in system.force("NonbondedForce").grid)
    grid = stated if all(stated) else o.default_grid(system.boxVectors)
    pos, vel = positions.copy(), velocities.copy()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.run_lfmiddle(pos, vel, system.boxVectors, 1, DT_PS, temperature, 1.0, 1e-6, seed, steps, grid)
        steps += 1
        if time.perf_counter() - t0 >= budget_s or steps >= 200:
            break
    el = time.perf_counter() - t0
    return system.particleCount * steps / el, steps, el


def _oracle_rate_job(job):
    """One worker's CPU baseline: the oracle on one replica for ~budget s."""
    system, positions, velocities, budget, temperature, seed = job
    rate, steps, el = oracle_rate(system, positions, velocities, budget, temperature, seed)
    return rate, steps, el, system.particleCount


def host_cpu():
    """(logical CPUs usable here, the CPU model name)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, model


def force_error(engine, system):
    """Relative RMS force error against the oracle at the engine's state:
    (counting the pairs within 3e-7 nm of the cutoff either way, as the
    parity tests do -- f32 and f64 may decide them differently -- and plain)."""
    import edge
    import oracle as O
    x = engine.positions()
    _, _, f = engine._native.evaluate(True)
    o = O.OracleSystem(system)
    _, _, of = o.potential(x, system.boxVectors, grid=engine.mesh)
    plain = float(np.sqrt(np.sum((f - of) ** 2) / np.sum(of ** 2)))
    aware, _ = edge.rel_rms_edge(f, of, system, x, system.boxVectors, o.alpha())
    return float(aware), plain


def health(engines, reps, atoms, when):
    """Log the replicas whose constraints no longer hold (a diverged run)."""
    bad = [(atoms[k], f"{d:.1e}") for k, (e, r) in enumerate(zip(engines, reps))
           for d in [constraint_error(e, r.built.system)] if not d < 1e-4]
    if bad:
        log(f"WARNING {when}: replicas whose constraints no longer hold (atoms, deviation): {bad}")
    else:
        log(f"{when}: every replica's constraints hold")
    return not bad


def shake_rate(engines, before):
    """SHAKE passes per cluster-step over `engines` since the counts `before`."""
    from mdcore import native
    passes = clusters = 0
    for e, (p0, c0) in zip(engines, before):
        p1, c1 = native.shake_counts(e._native)
        passes += p1 - p0
        clusters += c1 - c0
    return passes / clusters if clusters else 0.0


def constraint_error(engine, system):
    """Largest relative deviation of a constrained distance from its length."""
    if system.constraintCount == 0:
        return 0.0
    x = engine.positions().reshape(-1, 3)
    d = np.linalg.norm(x[system.constraintP1] - x[system.constraintP2], axis=1)
    return float(np.max(np.abs(d - system.constraintDistance) / system.constraintDistance))


def pair_roofline(pairs, kernel_ms, traffic=None, traffic_source=None):
    achieved = pairs * PAIR_FLOPS / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    return {
        "kernel": "k_nb_pairs (LJ + erfc-Ewald cluster-pair kernel)",
        "bound": "valu",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "traffic": traffic,
        "traffic_source": traffic_source,
        "algorithmic": f"{PAIR_FLOPS} flop x {int(pairs)} in-cutoff pairs per launch",
        "kernel_ms": kernel_ms,
    }


def run_ensemble(args):
    from mdcore.engine import HipBatch, HipEngine, RunOptions, minimise
    from mdcore.state import OpenMMState

    t = time.perf_counter()
    workers = max(1, min(16, os.cpu_count() or 1))
    if args.ensemble_cache and os.path.exists(args.ensemble_cache):
        # replicas this script built and saved earlier (same plan and seeds)
        import pickle
        with open(args.ensemble_cache, "rb") as f:
            reps = pickle.load(f)
        workers = 0
    else:
        reps = build_ensemble(args.replicas, workers)
        if args.ensemble_cache:
            import pickle
            with open(args.ensemble_cache, "wb") as f:
                pickle.dump(reps, f)
    atoms = [r.built.system.particleCount for r in reps]
    total_atoms = int(sum(atoms))
    log(f"built {len(reps)} replicas, {total_atoms} atoms ({min(atoms)}-{max(atoms)}) in "
        f"{time.perf_counter() - t:.1f} s with {workers} workers")
    if args.build_only:
        return None

    t = time.perf_counter()
    engines = []
    gave_up = 0
    for r in reps:
        # minimise, then equilibrate gently (0.5 fs, 10/ps friction): straight
        # from the minimiser to 2 fs, strained spots of a generated chain can
        # blow up within tens of steps; the reference protocol equilibrates
        # between minimisation and production for the same reason
        opts = dict(temperature=r.temperature, constraintTolerance=1e-6, seed=r.seed, maxBatch=args.max_batch)
        if args.verlet_skin is not None:
            opts["verletSkin"] = args.verlet_skin
        g = HipEngine(r.built.system, r.built.state,
                      RunOptions(timeStep=args.equilibrate_dt, friction=args.equilibrate_friction, **opts))
        res = minimise(g, args.minimise_iterations, 10.0)
        gave_up += int(res.gaveUp)
        g.setVelocitiesToTemperature(r.temperature, r.seed)
        g.run(args.equilibrate_steps)
        st = OpenMMState(boxVectors=r.built.state.boxVectors, positions=g.positions(), velocities=g.velocities())
        g.destroy()
        e = HipEngine(r.built.system, st, RunOptions(timeStep=DT_PS, friction=1.0, **opts))
        e.run(args.warmup)
        engines.append(e)
        if len(engines) % 8 == 0:
            log(f"  {len(engines)} replicas ready ({time.perf_counter() - t:.0f} s)")
    device_synchronize()
    log(f"minimised, equilibrated and warmed up {len(engines)} replicas in {time.perf_counter() - t:.1f} s "
        f"({gave_up} minimisations gave up)")
    health(engines, reps, atoms, "after warm-up")

    # ---- plain: replicas one after another through the single-system path --
    plain_steps = args.plain_steps
    for e in engines:  # capture each replica's graphs outside the timed region
        e.run(50)
        if plain_steps % 50:
            e.run(plain_steps % 50)
    device_synchronize()
    t0 = time.perf_counter()
    for e in engines:
        e.run(plain_steps)
    device_synchronize()
    plain_s = time.perf_counter() - t0
    plain_nsday = len(engines) * plain_steps * DT_PS * 1e-3 / plain_s * 86400.0
    log(f"plain: {len(engines)} x {plain_steps} steps in {plain_s:.2f} s = {plain_nsday:.1f} ns/day aggregate")
    health(engines, reps, atoms, "after the plain run")

    # ---- batched, plain SHAKE ----------------------------------------------------
    from mdcore import native
    batch = HipBatch(engines)
    # the pair kernel's schedule for this batch, picked by timing (changes no bits)
    tune = batch.autotune(args.tune_steps) if args.tune_steps > 0 else None
    if args.pair_schedule != "tuned" or args.pair_segment > 0:
        batch._native.set_pair_schedule(segment_entries=args.pair_segment,
                                        compact=args.pair_schedule == "compact" or
                                        (args.pair_schedule == "tuned" and bool(tune and tune["compact_pairs"])))
        log(f"pair schedule set to {args.pair_schedule}, {args.pair_segment or 'unchanged'} entries per unit")
    if tune:
        log(f"autotuned pair schedule: {tune['segment_entries']} entries per unit, {tune['pair_waves']} waves "
            f"({tune['ms_per_step']:.3f} ms/step phase by phase) vs the default {tune['default_segment_entries']}, "
            f"{tune['default_pair_waves']} ({tune['default_ms_per_step']:.3f} ms/step)")
    # the work-unit length the timed run uses (its measured traffic, if any)
    unit_entries = args.pair_segment or (tune["segment_entries"] if tune else None)
    batch.run(max(args.warmup, 50))
    batch.run(args.steps)  # captures the graph of the timed run's length outside the timed region
    device_synchronize()
    counts0 = [native.shake_counts(e._native) for e in engines]
    t0 = time.perf_counter()
    batch.run(args.steps)
    device_synchronize()
    cold_s = time.perf_counter() - t0
    cold_rate = shake_rate(engines, counts0)
    cold_nsday = len(engines) * args.steps * DT_PS * 1e-3 / cold_s * 86400.0
    log(f"batched, plain SHAKE: {args.steps} steps of all replicas in {cold_s:.2f} s = "
        f"{cold_nsday:.1f} ns/day aggregate, {cold_nsday / plain_nsday:.2f}x plain, "
        f"{cold_rate:.3f} SHAKE passes per cluster-step")
    health(engines, reps, atoms, "after the batched run")

    # ---- the learned SHAKE warm start: record, train, attach ---------------------
    t = time.perf_counter()
    ws = native.NativeWarmStart(seed=args.seed, capacity=args.ws_capacity)
    for e in engines:
        ws.record(e._native, args.record_steps)
    health(engines, reps, atoms, "after recording")
    loss0 = ws.train(1, args.train_batch, args.learning_rate)
    loss1 = ws.train(args.train_iterations, args.train_batch, args.learning_rate)
    for e in engines:
        ws.attach(e._native)
    device_synchronize()
    train_s = time.perf_counter() - t
    log(f"warm start: {ws.samples} samples from {args.record_steps} steps of every replica, "
        f"{args.train_iterations} Adam steps (loss {loss0:.4g} -> {loss1:.4g}) in {train_s:.1f} s")

    # ---- batched + learned: the measured configuration ---------------------------
    batch.run(max(args.warmup, 50))
    batch.run(args.steps)  # the cluster kernel changed: capture the timed length again
    device_synchronize()
    counts0 = [native.shake_counts(e._native) for e in engines]
    rebuilds0 = [e._native.list_flags()["rebuilds"] for e in engines]
    t0 = time.perf_counter()
    batch.run(args.steps)
    device_synchronize()
    batch_s = time.perf_counter() - t0
    rebuilds = sum(e._native.list_flags()["rebuilds"] - r0 for e, r0 in zip(engines, rebuilds0))
    steps_per_rebuild = len(engines) * args.steps / max(1, rebuilds)
    log(f"list rebuilds in the timed run: {rebuilds} over {len(engines)} replicas "
        f"(one per {steps_per_rebuild:.1f} replica-steps)")
    warm_rate = shake_rate(engines, counts0)
    batch_nsday = len(engines) * args.steps * DT_PS * 1e-3 / batch_s * 86400.0
    ms_per_step = batch_s / args.steps * 1e3
    log(f"batched + learned: {args.steps} steps of all replicas in {batch_s:.2f} s = {ms_per_step:.3f} ms/step, "
        f"{batch_nsday:.1f} ns/day aggregate, {batch_nsday / plain_nsday:.2f}x plain, "
        f"{warm_rate:.3f} SHAKE passes per cluster-step")
    deviations = [constraint_error(e, r.built.system) for e, r in zip(engines, reps)]
    worst_constraint = max(deviations)
    diverged = sum(1 for d in deviations if not d < 1e-4)
    log(f"worst relative constraint deviation after the learned run: {worst_constraint:.2e}")
    if worst_constraint > 1e-4:
        bad = [(atoms[k], f"{d:.1e}") for k, d in enumerate(deviations) if d > 1e-4]
        log(f"WARNING: replicas whose constraints no longer hold (atoms, deviation): {bad}")

    # ---- fixed force error: sampled replicas against the oracle ---------------
    order = np.argsort(atoms)
    sample = [int(order[0]), int(order[len(order) // 2])]
    both = {int(atoms[k]): force_error(engines[k], reps[k].built.system) for k in sample}
    errors = {k: v[0] for k, v in both.items()}
    errors_plain = {k: v[1] for k, v in both.items()}
    log(f"force rel RMS vs oracle after the batched run: {errors} (pairs at the cutoff either way), "
        f"{errors_plain} (plain)")

    # ---- roofline of the dominant kernel, measured inside the batched step -----
    phases = batch.profileStep(args.profile_steps)
    pairs = 0
    for e in engines:
        p, _ = e._native.pair_census()
        pairs += p
    log("batched phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()))
    integ_gbs = 124.0 * total_atoms / (phases["integrate"] * 1e-3) / 1e9 if phases["integrate"] > 0 else 0.0
    meshes = [tuple(e.mesh) for e in engines]
    points = float(sum(m[0] * m[1] * m[2] for m in meshes))

    def hbm(phase, nbytes):
        ms = phases.get(phase, 0.0)
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "bytes": nbytes, "ms": ms}

    # algorithmic bytes per launch, per unit of work: spreading 28 B/atom + 8 B
    # per grid point; the five transform passes (C2C forward + inverse, 3 axes
    # each) 96 B/point plus the convolution's 16; gathering 28 B/atom + 8 B per
    # point + 24 B/atom of force
    pme_roofline = {
        "pme_spread": hbm("pme_spread", 28.0 * total_atoms + 8.0 * points),
        "pme_fft_convolve": hbm("pme_fft_convolve", (96.0 + 16.0) * points),
        "pme_gather": hbm("pme_gather", (28.0 + 24.0) * total_atoms + 8.0 * points),
        "grid_points": points,
    }

    result = {
        "metric": METRIC,
        # a run in which any replica's constraints stopped holding does not stand
        "value": batch_nsday if diverged == 0 else None,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator: chains + TIP3P + ions), minimised, equilibrated 400 x 0.5 fs "
                "at 10/ps friction, then warmed up at 2 fs",
        "config": {
            "workload": f"batched ensemble + learned SHAKE warm start: {len(engines)} replicas, {total_atoms} atoms "
                        f"(log-uniform {min(atoms)}-{max(atoms)}), cubic/octahedral cells, 290-320 K, PME 1.0 nm "
                        "with 32^3-64^3 meshes by size, LFMiddle, dt 2 fs, SHAKE + SETTLE; one launch per phase "
                        "for all replicas, steps replayed from captured HIP graphs (50-step strides, shorter runs "
                        "their own length)",
            "replicas": len(engines),
            "atoms": total_atoms,
            "plain_ns_day": plain_nsday,
            "plain_seconds": plain_s,
            "plain_steps_per_replica": plain_steps,
            "batched_plain_shake_ns_day": cold_nsday,
            "batched_learned_over_plain": batch_nsday / plain_nsday,
            "batched_over_plain": cold_nsday / plain_nsday,
            "learned_over_batched": batch_nsday / cold_nsday,
            "atom_steps_per_s": total_atoms * args.steps / batch_s,
            "shake_passes_per_cluster_step": {"plain": cold_rate, "learned": warm_rate},
            "learned_warm_start": {"samples": int(ws.samples), "record_steps_per_replica": args.record_steps,
                                   "adam_steps": args.train_iterations, "batch": args.train_batch,
                                   "loss_first": loss0, "loss_last": loss1, "seconds": train_s,
                                   "network": "32-64-64-16 tanh MLP, bf16 MFMA 16x16x16, fp32 accumulation"},
            "worst_constraint_rel_deviation": worst_constraint,
            "diverged_replicas": diverged,
            "steps_per_list_rebuild": steps_per_rebuild,
            "autotune": tune,
            "force_rel_rms_vs_oracle": errors,
            "force_rel_rms_vs_oracle_plain": errors_plain,
            "in_cutoff_pairs": int(pairs),
            "arithmetic": "fp32 with Q32.32 int64 accumulation",
            "parallelism": "single GPU",
        },
        "roofline": pair_roofline(pairs, phases["pairs"], PAIR_TRAFFIC_BYTES.get(unit_entries),
                                  f"profiles/k_nb_pairs_ensemble64_traffic.md ({unit_entries}-entry units; rocprofv3 "
                                  "FETCH_SIZE + WRITE_SIZE, scaled by a 256 MiB copy)"),
        "phases_ms": phases,
        "integrator_roofline": {"bound": "hbm", "achieved": integ_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": integ_gbs / HBM_PEAK_GBS, "bytes_per_atom": 124},
        "pme_roofline": pme_roofline,
    }
    result["config"]["meshes"] = {f"{n}^3": sum(1 for m in meshes if m == (n, n, n))
                                  for n in sorted({m[0] for m in meshes})}
    if diverged:
        result["error"] = f"{diverged} replica(s) no longer hold their constraints: the figure does not stand"
    if not args.no_cpu_baseline:
        # The reference CPU path is single-threaded; the ensemble's replicas are
        # independent, so the host's cores run one replica each: `workers`
        # oracle processes on replicas spread over the size range, each timed
        # for ~cpu_seconds, their atom-step rates summed.  One core alone is
        # reported beside it.
        usable, model = host_cpu()
        workers = max(1, min(args.cpu_workers, usable))
        picks = [int(order[round(i * (len(order) - 1) / max(1, workers - 1))]) for i in range(workers)]
        log(f"timing the CPU oracle: {workers} processes on {usable} usable cores ({model}), "
            f"~{args.cpu_seconds:.0f} s each")
        jobs = [(reps[k].built.system, engines[k].positions(), engines[k].velocities(), args.cpu_seconds,
                 reps[k].temperature, reps[k].seed) for k in picks]
        if workers > 1:
            import multiprocessing as mp
            with mp.get_context("spawn").Pool(workers) as pool:
                timed = pool.map(_oracle_rate_job, jobs, chunksize=1)
        else:
            timed = [_oracle_rate_job(j) for j in jobs]
        all_rate = float(sum(r for r, _, _, _ in timed))
        one_rate = float(timed[0][0])  # the smallest replica on one core
        # the whole ensemble advancing one step costs total_atoms atom-steps
        cpu_nsday = len(engines) * DT_PS * 1e-3 * all_rate / total_atoms * 86400.0
        one_nsday = len(engines) * DT_PS * 1e-3 * one_rate / total_atoms * 86400.0
        result["cpu_baseline"] = {
            "value": cpu_nsday, "unit": "ns/day", "cores": workers, "kind": "port",
            "sample": f"the f64 C oracle (LFMiddle with SETTLE/SHAKE and the thermostat) in {workers} processes, "
                      f"one replica each ({min(a for _, _, _, a in timed)}-{max(a for _, _, _, a in timed)} atoms), "
                      f"~{args.cpu_seconds:.0f} s each; their atom-step rates summed and scaled to the whole "
                      "ensemble",
            "atom_steps_per_s": all_rate, "host_cpus": usable, "cpu_model": model,
            "single_core": {"value": one_nsday, "unit": "ns/day", "atom_steps_per_s": one_rate,
                            "replica_atoms": timed[0][3]},
            "per_process": [{"atoms": a, "steps": st, "seconds": el, "atom_steps_per_s": r}
                            for r, st, el, a in timed]}
        log(f"cpu baseline {cpu_nsday:.4f} ns/day aggregate on {workers} cores, {one_nsday:.5f} on one")
    batch.destroy()
    for e in engines:
        native.NativeWarmStart.detach(e._native)
        e.destroy()
    ws.close()
    return result


def run_traffic_probe(args):
    """A short run for hardware-counter passes (rocprofv3 --pmc): the ensemble's
    replicas as built (no minimisation or equilibration, 0.5 fs steps at 10/ps
    friction so the raw geometry stays bounded), a few batched steps, then one
    device-to-device copy of COPY_BYTES whose counter reading scales the
    counter to bytes."""
    import ctypes

    from mdcore.engine import HipBatch, HipEngine, RunOptions

    if args.ensemble_cache and os.path.exists(args.ensemble_cache):
        import pickle
        with open(args.ensemble_cache, "rb") as f:
            reps = pickle.load(f)
    else:
        reps = build_ensemble(args.replicas, max(1, min(16, os.cpu_count() or 1)))
    engines = [HipEngine(r.built.system, r.built.state,
                         RunOptions(timeStep=0.0005, friction=10.0, temperature=r.temperature,
                                    constraintTolerance=1e-6, seed=r.seed, maxBatch=1)) for r in reps]
    batch = HipBatch(engines)
    if args.pair_segment > 0:
        batch._native.set_pair_schedule(segment_entries=args.pair_segment)
    batch.run(args.traffic_probe)
    device_synchronize()
    log(f"traffic probe: {args.traffic_probe} batched steps of {len(engines)} replicas")
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(COPY_BYTES)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, ctypes.c_size_t(COPY_BYTES)) != 0:
        raise RuntimeError("hipMemset failed")
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(COPY_BYTES), 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy failed")
    device_synchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"traffic probe: copied {COPY_BYTES} bytes device to device")
    batch.destroy()
    for e in engines:
        e.destroy()


COPY_BYTES = 1 << 28


def run_single(args):
    from mdcore import synthetic
    from mdcore.engine import HipEngine, RunOptions, minimise

    built = synthetic.config(1)
    n = built.system.particleCount
    engine = HipEngine(built.system, built.state, RunOptions(timeStep=DT_PS, temperature=300.0, friction=1.0,
                                                            constraintTolerance=1e-6, seed=2026,
                                                            maxBatch=args.max_batch))
    res = minimise(engine, args.minimise_iterations, 10.0)
    log(f"minimised to {res.energy:.1f} kJ/mol in {res.iterations} iterations")
    engine.setVelocitiesToTemperature(300.0, 2026)
    engine.run(args.warmup)
    device_synchronize()
    t0 = time.perf_counter()
    engine.run(args.steps)
    device_synchronize()
    elapsed = time.perf_counter() - t0
    ns_day = args.steps * DT_PS * 1e-3 / elapsed * 86400.0
    log(f"timed {args.steps} steps: {elapsed / args.steps * 1e6:.1f} us/step, {ns_day:.1f} ns/day")
    pairs, _ = engine._native.pair_census()
    phases = engine.profileStep(args.profile_steps)
    log("phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()))
    err, err_plain = force_error(engine, built.system)
    result = {
        "metric": METRIC, "value": ns_day, "unit": "ns/day", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded generator: chain + ligand + TIP3P + ions), minimised and thermalised",
        "config": {"workload": f"plain single-system path: {n} atoms, truncated octahedron, PME 1.0 nm, LFMiddle "
                               "300 K, dt 2 fs, SHAKE + SETTLE, 50-step HIP-graph strides",
                   "atoms": n, "atom_steps_per_s": n * args.steps / elapsed, "in_cutoff_pairs": int(pairs),
                   "force_rel_rms_vs_oracle": err, "force_rel_rms_vs_oracle_plain": err_plain,
                   "parallelism": "single GPU"},
        "roofline": pair_roofline(pairs, phases["pairs"]),
        "phases_ms": phases,
    }
    if not args.no_cpu_baseline:
        rate, steps, el = oracle_rate(built.system, engine.positions(), engine.velocities(), args.cpu_seconds)
        result["cpu_baseline"] = {"value": DT_PS * 1e-3 * rate / n * 86400.0, "unit": "ns/day", "cores": 1,
                                  "kind": "port",
                                  "sample": f"{steps} LFMiddle steps of the same system by the f64 C oracle on one "
                                            f"host core, {el:.1f} s"}
    engine.destroy()
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", choices=("ensemble", "single"), default="ensemble")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="oracle processes for the CPU baseline (at most the usable cores)")
    ap.add_argument("--minimise-iterations", type=int, default=60)
    ap.add_argument("--equilibrate-steps", type=int, default=400)
    ap.add_argument("--equilibrate-dt", type=float, default=0.0005, help="ps")
    ap.add_argument("--equilibrate-friction", type=float, default=10.0, help="1/ps")
    ap.add_argument("--record-steps", type=int, default=25, help="warm-start samples: steps recorded per replica")
    ap.add_argument("--train-iterations", type=int, default=1500)
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--learning-rate", type=float, default=1e-3)
    ap.add_argument("--ws-capacity", type=int, default=4_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tune-steps", type=int, default=5,
                    help="steps timed per pair-kernel schedule by the autotuner (0: no tuning)")
    ap.add_argument("--pair-schedule", choices=("tuned", "direct", "compact"), default="tuned",
                    help="the pair kernel's evaluation: as the autotuner picked, or forced")
    ap.add_argument("--pair-segment", type=int, default=0,
                    help="list entries per pair-kernel work unit, forced (0: tuned, or the default in a probe)")
    ap.add_argument("--verlet-skin", type=float, default=None,
                    help="list skin in nm (default: the engine's, 0.15)")
    ap.add_argument("--max-batch", type=int, default=0,
                    help="steps per captured HIP graph (0: 50; 1: launch every step directly)")
    ap.add_argument("--ensemble-cache", default="",
                    help="build the ensemble once into this file and load it on later runs (so a profiled run "
                         "starts no builder processes)")
    ap.add_argument("--build-only", action="store_true", help="build (and cache) the ensemble, then stop")
    ap.add_argument("--traffic-probe", type=int, default=0, metavar="STEPS",
                    help="instead of the benchmark: STEPS batched steps of the ensemble and one copy of known size, "
                         "for rocprofv3 --pmc passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build is single-GPU (asked for --gpus {args.gpus}). "
              "Independent replicas would shard across GPUs with no communication between them.",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe > 0:
        run_traffic_probe(args)
        return
    from mdcore import native
    try:
        result = run_ensemble(args) if args.workload == "ensemble" else run_single(args)
    except native.MdError as err:
        # a replica diverged, overflowed its list or hit another engine error:
        # no figure stands, and the line says why
        print(json.dumps({"metric": METRIC, "value": None, "unit": "ns/day", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "fp32", "data": "synthetic", "config": {"workload": args.workload},
                          "error": str(err)}), flush=True)
        sys.exit(1)
    if result is not None:
        print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
