"""Benchmark of the force-and-integrate step on one MI355X.

Headline: direct-sum Ginteractions/s on the 50,000-particle exponential disk
(scale 20, radius 100; G=1, softening 0.2, dt=1e-3).  The same JSON line also
carries, in `secondary`, the Barnes-Hut step at 500,000 particles (two
exponential disks, theta=1, W=32: the group width of the reference's
published runs; --bh-width 64 for wave-wide groups), the batched 256-system
mix and the served request stream (Poisson 50/s for 60 s), because the
headline metric names all of them.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bh-steps K2] [--no-cpu]
                    [--batch-steps K3] [--no-batch] [--serve-seconds S]

Inputs are generated on the host and uploaded before the timed region; the
timed region covers K steps bracketed by device synchronisation.  Kernel times
come from HIP events recorded by libgsim on the simulation's own stream.  The
CPU leg times the oracle on the same inputs and, as the checker, states the
force error of the benchmarked GPU results (FP32 oracle and FP64 direct sum).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gravity-sim_amd", "python"))

import numpy as np  # noqa: E402

import gsim  # noqa: E402
from gsim import workloads as WL  # noqa: E402

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3          # vector FP32 (packed FMA), TFLOP/s
PEAK_HBM_GBS = 8000.0             # HBM3E, GB/s
VALU_LANE_OPS = 256 * 64 * 2.4e9  # one FP32 op per lane per clock, lane-ops/s
DS_FLOP_PER_PAIR = 14             # 2 sub, 2 FMA, rsqrt, 3 mul, 2 FMA (one-sided interaction)
DS_LANEOPS_PER_PAIR = 10
# symmetric kernel: per unordered pair 12 VALU + 1 rsq, plus 5 DPP rotates per
# 16 pairs (counted in its ISA); B(B+1)/2 block pairs of 64 x 64 (diagonal
# blocks one-sided)
DS_SYM_ISSUED_PER_PAIR = 13.3125
# HBM bytes per ds_sym_kernel launch at 50,048 particles: rocprofv3 FETCH_SIZE
# x 2 + WRITE_SIZE x 1 (factors from a 256 MiB copy), 2 x 2,417 + 48,950 KiB;
# profiles/ds50k_bh500k_counters.md
DS_TRAFFIC_BYTES = (2 * 2417 + 48950) * 1024
# HBM bytes per traversal launch at 500,032 particles (same counters, median
# over the steady steps, state in sorted order), by group width:
# bh_traverse64_kernel (one wave per group, staged runs) 2 x 44,890 + 25,530
# KiB; bh_traverse32_kernel (one 32-particle group per wave) 2 x 44,310 +
# 25,060 KiB
BH_TRAFFIC_BYTES = {64: (2 * 44890 + 25530) * 1024, 32: (2 * 44310 + 25060) * 1024}
BH_FLOP_VISIT, BH_FLOP_ACCEPT = 7, 8     # 2 sub + 2 FMA + theta-mul; rsqrt + 3 mul + 2 FMA
BH_LANEOPS_VISIT, BH_LANEOPS_ACCEPT = 6, 6
REF_DS_MS_50K = 3.602             # reference README.md:18 (RTX 3070), ms/step at 50,000
REF_BH_MS_500K = 2.853            # reference README.md:19, ms/step at 500,000

DS_N = WL.DS_N
BH_N = WL.BH_N
ERR_SAMPLE = 2048                 # targets of the FP64 error sample
CPU_THREADS_MAX = 16              # the GPU box's host share for one GPU


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _public(d: dict) -> dict:
    return {k: v for k, v in d.items() if not k.startswith("_")}


def bench_direct_sum(steps: int, warmup: int) -> dict:
    p = WL.direct_sum_disk()
    sim = gsim.Simulation(p, "direct-sum")
    sim.step(max(warmup, 1))
    sim.synchronize()
    sim.set_timing(True)
    sim.timing(reset=True)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    t = sim.timing(reset=True)
    sim.close()
    # PCIe-inclusive: host arrays in, one step, host arrays out; the same
    # step's accelerations are kept for the error statement
    t0 = time.perf_counter()
    with gsim.Simulation(p, "direct-sum") as once:
        once.step(1)
        once.read()
    pcie_ms = (time.perf_counter() - t0) * 1e3
    with gsim.Simulation(p, "direct-sum") as dbg:
        dbg.set_debug(gsim.DEBUG_ACCEL)
        dbg.step(1)
        accel = dbg.dump(gsim.DUMP_ACCEL)[: p.n].copy()
    pairs = float(DS_N) * DS_N
    ms_step = elapsed * 1e3 / steps
    kernel_ms = t["force_ms"] / max(t["force_launches"], 1)
    achieved = DS_FLOP_PER_PAIR * pairs / (kernel_ms * 1e-3) / 1e12
    blocks = (DS_N + 63) // 64
    issued = DS_SYM_ISSUED_PER_PAIR * blocks * (blocks + 1) / 2 * 4096
    return {
        "value": pairs * steps / elapsed / 1e9,
        "ms_per_step": ms_step,
        "kernel_ms": kernel_ms,
        "combine_ms": t["integrate_ms"] / max(t["steps"], 1),
        "pcie_inclusive_ms": pcie_ms,
        "roofline": {
            "bound": "valu",
            "kernel": "ds_sym_kernel",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": DS_TRAFFIC_BYTES,
            "note": "achieved counts the one-sided algorithm's 14 FLOP x N^2 per step over the "
                    "force kernel's mean HIP-event duration; the symmetric kernel evaluates each "
                    "unordered pair once; peak is the packed-FMA vector figure",
            "issued_valu_lane_op_frac": issued / (kernel_ms * 1e-3) / VALU_LANE_OPS,
        },
        "_input": p,
        "_accel": accel,
    }


def bench_barnes_hut(steps: int, warmup: int, leaf_bits: int = 10, box: bool = False,
                     width: int = 32) -> dict:
    """leaf_bits = 15 / box: the opt-in tree options beyond the reference
    (their trees differ from the reference's, so vs_baseline is null).
    width: the traversal group width W (32, the published runs', or 64).

    Two passes over the same K steps from the same post-warmup state: the
    timed pass (asynchronous, HIP events per phase) gives the step and
    kernel times; the counted pass replays the identical steps one at a time
    and reads each step's node visits and acceptances, so the roofline's
    work and time cover the same K steps."""
    p = WL.two_disks().padded()  # padded: the whole state reads back
    reference_tree = leaf_bits == 10 and not box

    def make(particles):
        s = gsim.Simulation(particles, "barnes-hut", accuracy_parameter=1.0, group_width=width)
        if leaf_bits != 10:
            s.set_tuning(gsim.TUNE_BH_LEAF_BITS, leaf_bits)
        if box:
            s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
        return s

    with make(p) as dbg:  # the first step's accelerations, for the error statement
        dbg.set_debug(gsim.DEBUG_ACCEL)
        dbg.step(1)
        accel = dbg.dump(gsim.DUMP_ACCEL).copy()
    sim = make(p)
    sim.step(max(warmup, 1))
    start = sim.read()
    sim.set_timing(True)
    sim.timing(reset=True)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    t = sim.timing(reset=True)
    sim.close()
    V = A = cells = 0
    with make(start) as cnt:
        for _ in range(steps):
            cnt.step(1)
            v, a = (int(x) for x in cnt.dump(gsim.DUMP_COUNTERS))
            V += v
            A += a
            cells += int(cnt.dump(gsim.DUMP_NODE_TOTAL)[0]) - p.n
    ms_step = elapsed * 1e3 / steps
    trav_ms = t["force_ms"] / max(t["steps"], 1)
    build_ms = t["build_ms"] / max(t["steps"], 1)
    flops = BH_FLOP_VISIT * V + BH_FLOP_ACCEPT * A
    achieved = flops / (t["force_ms"] * 1e-3) / 1e12
    lane_ops = BH_LANEOPS_VISIT * V + BH_LANEOPS_ACCEPT * A
    # step roofline over the K steps: build bytes at HBM peak + traversal
    # lane-ops at VALU peak.  Build bytes 64 per particle (keys 12, one sort
    # pass 16, leaf emission 36) + 20 per cell node (nodes - particles: none
    # leaves the domain here)
    build_bytes = 64 * p.n * steps + 20 * max(cells, 0)
    t_star = build_bytes / (PEAK_HBM_GBS * 1e9) + lane_ops / VALU_LANE_OPS
    return {
        "metric": "BH step ms @500k",
        "value": ms_step,
        "unit": "ms/step",
        "higher_is_better": False,
        "vs_baseline": REF_BH_MS_500K / ms_step if reference_tree else None,
        "particle_steps_per_s": BH_N / (ms_step * 1e-3),
        "build_ms": build_ms,
        "traverse_ms": trav_ms,
        "visits_per_step": V / steps,
        "accepts_per_step": A / steps,
        "config": {"workload": "barnes-hut, two exponential disks of 250,000 (radius 33, "
                               f"scale 7, centres +-40, bulk vx -+2), theta=1, W={width}, "
                               "G=1, softening 0.2, dt=1e-3", "particles": BH_N,
                   "theta": 1.0, "group_width": width, "leaf_bits": leaf_bits,
                   "adaptive_box": box},
        "roofline": {
            "bound": "valu",
            "kernel": f"bh_traverse{width}_kernel",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": BH_TRAFFIC_BYTES[width] if leaf_bits == 10 and not box else None,
            "valu_lane_op_frac": lane_ops / (t["force_ms"] * 1e-3) / VALU_LANE_OPS,
            "build_bytes_per_step": build_bytes / steps,
            "build_hbm_frac": build_bytes / (t["build_ms"] * 1e-3) / (PEAK_HBM_GBS * 1e9),
            "step_frac": t_star / (elapsed),
            "note": "work (7V + 8A FLOP, 6V + 6A lane-ops) and time cover the same K steps; "
                    "step_frac = (build bytes / HBM peak + lane-ops / VALU peak) / wall time",
        },
        "_input": p,
        "_accel": accel,
    }


def bench_batched(steps: int, serial_steps: int = 20) -> dict:
    systems = WL.batched_systems()
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    batch = gsim.Batch(cap, max_systems=len(systems))
    for p, algo, kw in systems:
        batch.add(p, algo, **kw)
    t0 = time.perf_counter()
    batch.autotune(steps_per_trial=5)
    tune_s = time.perf_counter() - t0
    picks = batch.tuning()
    batch.step(2)
    batch.synchronize()
    v0, a0 = batch.visits(), batch.accepts()
    t0 = time.perf_counter()
    batch.step(steps)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
    V, A = batch.visits() - v0, batch.accepts() - a0
    batch.close()
    # the same systems one by one through the plain single-system path
    sims = [gsim.Simulation(p, algo, **kw) for p, algo, kw in systems]
    for s in sims:
        s.step(1)
    for s in sims:
        s.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.step(serial_steps)
        s.synchronize()
    t_serial = time.perf_counter() - t0
    for s in sims:
        s.close()
    n_ds = sum(1 for _, a, _ in systems if a == "direct-sum")
    ds_pairs = sum(((p.n + 63) // 64 * 64) ** 2 for p, a, _ in systems if a == "direct-sum")
    lane_ops = (BH_LANEOPS_VISIT * V + BH_LANEOPS_ACCEPT * A
                + DS_LANEOPS_PER_PAIR * ds_pairs * steps)
    batched = len(systems) * steps / t_batch
    serial = len(systems) * serial_steps / t_serial
    return {
        "metric": "batched system-steps/s (256 mixed systems)",
        "value": batched,
        "unit": "system-steps/s",
        "higher_is_better": True,
        "vs_baseline": None,
        "ms_per_step": t_batch * 1e3 / steps,
        "one_by_one_value": serial,
        "speedup_vs_one_by_one": batched / serial,
        "particles": int(sum(p.n for p, _, _ in systems)),
        "interactions_per_s": (A + ds_pairs * steps) / t_batch,
        "roofline": {
            "bound": "valu",
            "achieved_lane_ops_per_s": lane_ops / t_batch,
            "peak": VALU_LANE_OPS,
            "unit": "lane-ops/s",
            "frac": lane_ops / t_batch / VALU_LANE_OPS,
            "note": "6 lane-ops per Barnes-Hut visit and per acceptance, 10 per direct-sum "
                    "pair, over the whole step's wall time (the build's HBM bytes excluded)",
        },
        "config": {"workload": f"{len(systems)} systems, N log-uniform [1k, 64k], random or "
                               f"exponential disk (radius 100), N < 8192 direct sum ({n_ds}), "
                               f"else Barnes-Hut theta=0.7 ({len(systems) - n_ds}); one hipGraph "
                               f"replay per step", "steps": steps,
                   "one_by_one_steps": serial_steps,
                   "autotuned": dict(picks, seconds=tune_s,
                                     method="coordinate descent over waves per SIMD, sort "
                                            "tile, run queue, priority step and carry; each "
                                            "candidate 2 x 5 steps from a restored snapshot")},
    }


def bench_serve(seconds: float, budget_gb: float = 200.0, max_active: int = 0,
                deadline_s: float = 420.0, one_by_one_sample: int = 24) -> dict:
    """The served stream: requests submitted at their Poisson arrival times,
    drained until idle or `deadline_s` after the first arrival.  Then a
    seeded random sample of the same requests is run alone, one after
    another, through the single-system path: its particle-steps/s, applied
    to the whole stream's work, gives the one-by-one systems/s."""
    reqs = WL.served_requests(seconds)
    offered = sum(r.particles.n * r.steps for r in reqs)
    log(f"  {len(reqs)} requests, {offered / 1e9:.1f} G particle-steps offered over {seconds:g} s")
    srv = gsim.Server(int(budget_gb * 1e9), max_systems=2048, max_active_particles=max_active)
    rids = []

    def client():
        t0 = time.perf_counter()
        for r in reqs:
            wait = r.at - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            rids.append(srv.submit(r.particles, r.steps, "barnes-hut",
                                   accuracy_parameter=r.theta))

    th = threading.Thread(target=client)
    t0 = time.perf_counter()
    th.start()
    while th.is_alive():  # arrivals: a progress line every 20 s of them
        th.join(20.0)
        if th.is_alive():
            st = srv.stats()
            log(f"  arrivals: {len(rids)} submitted, {st['completed']} done")
    # drain in slices, reporting what has completed, until idle or the deadline
    deadline = t0 + deadline_s
    drained = False
    while not drained and time.perf_counter() < deadline:
        drained = srv.drain(min(30.0, max(deadline - time.perf_counter(), 0.0)))
        if not drained:
            st = srv.stats()
            log(f"  draining: {st['completed']} of {len(rids)} done")
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat, queue, work_done = [], [], 0
    for rid, r in zip(rids, reqs):
        if srv.wait(rid, 0.0):
            _, info = srv.result(rid)
            lat.append(info["latency_ms"])
            queue.append(info["queue_ms"])
            work_done += r.particles.n * r.steps
    srv.close()
    span = max(st["span_s"], 1e-9)
    inter = st["bh_interactions"] + st["ds_interactions"]
    lane_ops = (BH_LANEOPS_VISIT * st["bh_visits"] + BH_LANEOPS_ACCEPT * st["bh_interactions"]
                + DS_LANEOPS_PER_PAIR * st["ds_interactions"])
    # one by one: a seeded random sample run alone
    rng = np.random.default_rng(17)
    pick = sorted(rng.choice(len(reqs), size=min(one_by_one_sample, len(reqs)), replace=False))
    sample_work, t_alone = 0, 0.0
    for k in pick:
        r = reqs[k]
        with gsim.Simulation(r.particles, "barnes-hut", accuracy_parameter=r.theta) as s:
            s.synchronize()
            ta = time.perf_counter()
            s.step(r.steps)
            s.synchronize()
            t_alone += time.perf_counter() - ta
        sample_work += r.particles.n * r.steps
    alone_rate = sample_work / t_alone
    one_by_one = len(reqs) / (offered / alone_rate)
    value = st["completed"] / span
    return {
        "metric": "served systems/s (continuous batching)",
        "value": value,
        "unit": "systems/s",
        "higher_is_better": True,
        "vs_baseline": None,
        "completed": st["completed"],
        "requests": len(reqs),
        "drained": drained,
        "p50_ms": float(np.percentile(lat, 50)) if lat else None,
        "p99_ms": float(np.percentile(lat, 99)) if lat else None,
        "queue_p50_ms": float(np.percentile(queue, 50)) if queue else None,
        "interactions_per_s": inter / span,
        "particle_steps_per_s": work_done / span,
        "offered_particle_steps_per_s": offered / seconds,
        "span_s": span,
        "wall_s": wall,
        "batch_steps": st["batch_steps"],
        "one_by_one_value": one_by_one,
        "speedup_vs_one_by_one": value / one_by_one,
        "one_by_one_particle_steps_per_s": alone_rate,
        "one_by_one_sample": f"{len(pick)} requests drawn at random (seed 17) from the stream, "
                             "each run alone for its steps through gsim_create/gsim_step, one "
                             "after another; their particle-steps/s applied to the stream's work",
        "roofline": {
            "bound": "valu",
            "achieved_lane_ops_per_s": lane_ops / span,
            "peak": VALU_LANE_OPS,
            "unit": "lane-ops/s",
            "frac": lane_ops / span / VALU_LANE_OPS,
            "note": "6 lane-ops per Barnes-Hut visit and per acceptance over the span from "
                    "the first arrival to the last completion (arrival gaps and the build's "
                    "HBM bytes included in the time)",
        },
        "config": {"workload": f"Poisson arrivals 50/s for {seconds:g} s ({len(reqs)} requests), "
                               "N log-uniform [2k, 2M], preset random/galaxy/galaxy-collision "
                               "(size 100, masses U[1,10]), Barnes-Hut theta in {0.5,0.7,1.0}, "
                               f"steps U{{10..1000}}, HBM budget {budget_gb:g} GB",
                   "capacity_particles": st["capacity_particles"], "arenas": st["arenas"],
                   "max_active_particles": max_active, "deadline_s": deadline_s},
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _err_stats(a, ref) -> dict:
    """Per-particle |a - ref| / |ref| (median, p99, max), the same relative to
    max(|ref|, rms |ref|), and the system-relative error."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    d = np.linalg.norm(a - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt(np.mean(mag ** 2)))
    rel = d / np.maximum(mag, 1e-300)
    relr = d / np.maximum(mag, rms)
    return {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
            "max": float(rel.max()), "rms_rel_max": float(relr.max()),
            "system": float(np.linalg.norm(a - ref) / np.linalg.norm(ref))}


def cpu_baseline(ds: dict, budget_s: float = 12.0) -> tuple[dict, dict]:
    """The oracle's FP32 direct sum (scalar C, one thread per core slice) on
    the same 50k workload, timed on this host: a bounded sample of whole
    force passes over all 50,000 x 50,000 pairs on up to 16 threads, and one
    pass on one thread.  As the checker it also states the error of the
    GPU's accelerations on the same input: against the FP32 oracle over all
    particles, and against the FP64 direct sum on a sample of targets."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline and checker

    threads = max(1, min(CPU_THREADS_MAX, os.cpu_count() or 1))
    p = ds["_input"]
    eps2 = oracle.square_f32(0.2)
    t0 = time.perf_counter()
    ref32 = oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=threads)
    one = time.perf_counter() - t0
    reps = max(1, min(8, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=threads)
    el = time.perf_counter() - t0
    pairs = float(DS_N) * DS_N * reps
    t0 = time.perf_counter()
    oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=1)
    el1 = time.perf_counter() - t0
    cpu = {"value": pairs / el / 1e9, "unit": "Ginteractions/s", "cores": threads,
           "kind": "port",
           "one_core_value": float(DS_N) * DS_N / el1 / 1e9,
           "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
           "sample": f"{reps + 1} full 50,000x50,000 FP32 force passes of the oracle "
                     f"({threads} threads) and one on 1 thread, integrate excluded"}
    a = ds["_accel"]
    rng = np.random.default_rng(23)
    tg = np.sort(rng.choice(p.n, size=ERR_SAMPLE, replace=False)).astype(np.uint32)
    ref64 = oracle.ds_accel_f64_at(p.pos, p.mass, eps2, tg, nthreads=threads)
    err = {"vs_fp32_oracle": _err_stats(a, ref32),
           "vs_fp64": _err_stats(a[tg], ref64),
           "fp32_oracle_vs_fp64": _err_stats(ref32[tg], ref64),
           "tolerance": "system-relative <= 1e-5 against the FP32 oracle",
           "sample": f"FP64 on {ERR_SAMPLE} random targets (seed 23) against all 50,000 "
                     "sources; FP32 oracle on all particles; first step of the input"}
    return cpu, err


def cpu_baseline_bh(bh: dict, steps: int = 3, width: int = 32) -> tuple[dict, dict]:
    """The oracle's whole Barnes-Hut step (dense-pyramid build, group walk,
    integrate) on the same 500k workload and group width, on up to 16
    threads: ms per step.  As the checker: the GPU's first-step accelerations
    against the oracle's walk (FP32) and, on a sample of targets, against the
    FP64 direct sum (the Barnes-Hut theta error)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline and checker

    threads = max(1, min(CPU_THREADS_MAX, os.cpu_count() or 1))
    p = bh["_input"]
    eps2 = oracle.square_f32(0.2)
    theta2 = oracle.square_f32(1.0)
    pos, vel, mass = p.pos, p.vel, p.mass
    pos, vel, mass, _, _ = oracle.bh_step(pos, vel, mass, 1.0, 1e-3, eps2, theta2, width, threads)
    t0 = time.perf_counter()
    for _ in range(steps):
        pos, vel, mass, _, _ = oracle.bh_step(pos, vel, mass, 1.0, 1e-3, eps2, theta2, width,
                                              threads)
    el = time.perf_counter() - t0
    cpu = {"value": el * 1e3 / steps, "unit": "ms/step", "cores": threads, "kind": "port",
           "sample": f"{steps + 1} whole Barnes-Hut steps of the oracle at {p.n:,} "
                     f"particles, W={width} ({threads} threads)"}
    tree, _, _, pk, mk = oracle.bh_build(p.pos, p.mass)
    oa = oracle.bh_traverse(tree, pk, mk, eps2, theta2, W=width)[0]
    a = bh["_accel"]
    rng = np.random.default_rng(29)
    tg = np.sort(rng.choice(BH_N, size=ERR_SAMPLE, replace=False)).astype(np.uint32)
    ref64 = oracle.ds_accel_f64_at(pk, mk, eps2, tg, nthreads=threads)
    err = {"vs_fp32_oracle": _err_stats(a[:BH_N], oa[:BH_N]),
           "theta_error_vs_fp64": _err_stats(a[tg], ref64),
           "fp32_oracle_theta_error_vs_fp64": _err_stats(oa[tg], ref64),
           "tolerance": "system-relative <= 1e-5 against the FP32 oracle's walk (same tree, "
                        "same visits and acceptances)",
           "sample": f"FP64 direct sum on {ERR_SAMPLE} random targets (seed 29) against all "
                     f"{BH_N:,} sources; FP32 oracle walk on all particles; first step"}
    return cpu, err


def copy_calibration(mib: int = 256, reps: int = 3) -> None:
    """Device-to-device copies of a known size, the reference point for
    scaling the FETCH_SIZE / WRITE_SIZE counters of the kernels above."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB x {reps}")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bh-steps", type=int, default=20)
    ap.add_argument("--bh-width", type=int, default=32, choices=(32, 64),
                    help="Barnes-Hut group width W (32: the reference's published runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bh", action="store_true")
    ap.add_argument("--bh-leaf-bits", type=int, default=10, choices=(10, 15),
                    help="Barnes-Hut leaf grid bits per axis (15: beyond the reference)")
    ap.add_argument("--bh-adaptive-box", action="store_true",
                    help="Barnes-Hut tree over the bounding square (beyond the reference)")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="arrival window of the served workload (0: skip)")
    ap.add_argument("--serve-deadline", type=float, default=420.0,
                    help="seconds after the first arrival at which the served run stops "
                         "draining and reports what completed")
    ap.add_argument("--serve-max-active", type=int, default=0,
                    help="cap on particles stepping at once in the served run (0: none)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run known-size device copies (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every GPU box of this build has one GPU and "
              "the simulation step has no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        return 2
    log(f"direct-sum 50k: warmup {args.warmup}, steps {args.steps}")
    ds = bench_direct_sum(args.steps, args.warmup)
    log(f"  {ds['value']:.1f} Ginteractions/s, {ds['ms_per_step']:.4f} ms/step")
    bh = None
    if not args.no_bh:
        log(f"barnes-hut 500k: steps {args.bh_steps}, {args.bh_leaf_bits}-bit leaves, "
            f"W={args.bh_width}" + (", adaptive box" if args.bh_adaptive_box else ""))
        bh = bench_barnes_hut(args.bh_steps, max(args.warmup, 1), args.bh_leaf_bits,
                              args.bh_adaptive_box, args.bh_width)
        log(f"  {bh['value']:.4f} ms/step (build {bh['build_ms']:.4f}, traverse "
            f"{bh['traverse_ms']:.4f}), step fraction {bh['roofline']['step_frac']:.3f}")
    bt = None
    if not args.no_batch:
        log(f"batched 256 systems: steps {args.batch_steps}")
        bt = bench_batched(args.batch_steps)
        log(f"  {bt['value']:.0f} system-steps/s ({bt['ms_per_step']:.3f} ms/step), "
            f"{bt['speedup_vs_one_by_one']:.2f}x one by one")
    sv = None
    if args.serve_seconds > 0:
        log(f"served: {args.serve_seconds:g} s of arrivals")
        sv = bench_serve(args.serve_seconds, max_active=args.serve_max_active,
                         deadline_s=args.serve_deadline)
        log(f"  {sv['value']:.2f} systems/s, p50 {sv['p50_ms']} ms, p99 {sv['p99_ms']} ms, "
            f"{sv['interactions_per_s'] / 1e9:.1f} Ginteractions/s, "
            f"{sv['speedup_vs_one_by_one']:.2f}x one by one")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu, ds["error"] = cpu_baseline(ds)
        log(f"  {cpu['value']:.2f} Ginteractions/s on {cpu['cores']} threads, "
            f"{cpu['one_core_value']:.3f} on one; force error vs FP32 oracle "
            f"{ds['error']['vs_fp32_oracle']['system']:.2e}")
        if bh is not None:
            bh["cpu_baseline"], bh["error"] = cpu_baseline_bh(bh, width=args.bh_width)
            log(f"  Barnes-Hut 500k: {bh['cpu_baseline']['value']:.1f} ms/step on "
                f"{bh['cpu_baseline']['cores']} threads; theta error vs FP64 median "
                f"{bh['error']['theta_error_vs_fp64']['median']:.2e}")
    ref_ginter = DS_N * DS_N / (REF_DS_MS_50K * 1e-3) / 1e9
    line = {
        "metric": "direct-sum Ginteractions/s @50k",
        "value": ds["value"],
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ds["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": ds["value"] / ref_ginter,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "direct-sum, 50,000-particle exponential disk (radius 100, "
                               "scale 20, circular velocities + 5% dispersion, masses "
                               "U[1,10]), G=1, softening 0.2, dt=1e-3",
                   "particles": DS_N, "parallelism": "single-gpu"},
        "pcie_inclusive_ms_one_step": ds["pcie_inclusive_ms"],
        "roofline": ds["roofline"],
        "cpu_baseline": cpu,
        "error": ds.get("error"),
        "secondary": [_public(x) for x in (bh, bt, sv) if x],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
