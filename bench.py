"""Benchmark of the force-and-integrate step on one MI355X.

Headline: direct-sum Ginteractions/s on the 50,000-particle exponential disk
(scale 20, radius 100; G=1, softening 0.2, dt=1e-3).  The same JSON line also
carries the Barnes-Hut step at 500,000 particles (two exponential disks,
theta=1, W=32: the group width of the reference's published runs; --bh-width
64 for wave-wide groups) because the headline metric names both.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bh-steps K2] [--no-cpu]
                    [--batch-steps K3] [--no-batch] [--serve-seconds S]

Inputs are generated on the host and uploaded before the timed region; the
timed region covers K steps bracketed by device synchronisation.  Kernel times
come from HIP events recorded by libgsim on the simulation's own stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gravity-sim_amd", "python"))

import numpy as np  # noqa: E402

import gsim  # noqa: E402

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3          # vector FP32 (packed FMA), TFLOP/s
PEAK_HBM_GBS = 8000.0             # HBM3E, GB/s
VALU_LANE_OPS = 256 * 64 * 2.4e9  # one FP32 op per lane per clock, lane-ops/s
DS_FLOP_PER_PAIR = 14             # 2 sub, 2 FMA, rsqrt, 3 mul, 2 FMA (one-sided interaction)
DS_LANEOPS_PER_PAIR = 10
# symmetric kernel: per unordered pair 12 VALU + 1 rsq, plus 5 DPP rotates per
# 16 pairs; B(B+1)/2 block pairs of 64 x 64 (diagonal blocks one-sided)
DS_SYM_ISSUED_PER_PAIR = 13.3125
# HBM bytes per ds_sym_kernel launch at 50,048 particles: rocprofv3 FETCH_SIZE
# x 2 + WRITE_SIZE x 1 (factors from a 256 MiB copy), 2 x 2,417 + 48,950 KiB;
# profiles/ds50k_bh500k_counters.md
DS_TRAFFIC_BYTES = (2 * 2417 + 48950) * 1024
# HBM bytes per traversal launch at 500,032 particles (same counters), by
# group width: bh_traverse64_kernel (one wave per group, staged runs)
# 2 x 75,874 + 33,931 KiB; bh_traverse32_kernel (one 32-particle group per
# wave) 2 x 82,905 + 33,739 KiB
BH_TRAFFIC_BYTES = {64: (2 * 75874 + 33931) * 1024, 32: (2 * 82905 + 33739) * 1024}
BH_FLOP_VISIT, BH_FLOP_ACCEPT = 7, 8     # 2 sub + 2 FMA + theta-mul; rsqrt + 3 mul + 2 FMA
BH_LANEOPS_VISIT, BH_LANEOPS_ACCEPT = 6, 6
REF_DS_MS_50K = 3.602             # reference README.md:18 (RTX 3070), ms/step at 50,000
REF_BH_MS_500K = 2.853            # reference README.md:19, ms/step at 500,000

DS_N = 50000
BH_N = 500000


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def bench_direct_sum(steps: int, warmup: int) -> dict:
    p = gsim.generate("exponential-disk", DS_N, 100.0, 1.0, 10.0, 1.0, seed=2)
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
    # PCIe-inclusive: host arrays in, one step, host arrays out
    t0 = time.perf_counter()
    with gsim.Simulation(p, "direct-sum") as once:
        once.step(1)
        once.read()
    pcie_ms = (time.perf_counter() - t0) * 1e3
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
            "note": "achieved counts the one-sided algorithm's 14 FLOP x N^2 per step; the "
                    "symmetric kernel evaluates each unordered pair once",
            "issued_valu_lane_op_frac": issued / (kernel_ms * 1e-3) / VALU_LANE_OPS,
        },
    }


def bench_barnes_hut(steps: int, warmup: int, leaf_bits: int = 10, box: bool = False,
                     width: int = 32) -> dict:
    """leaf_bits = 15 / box: the opt-in tree options beyond the reference
    (their trees differ from the reference's, so vs_baseline is null).
    width: the traversal group width W (32, the published runs', or 64)."""
    p = gsim.generate("two-exponential-disks", BH_N, 33.0, 1.0, 10.0, 1.0, seed=3)
    sim = gsim.Simulation(p, "barnes-hut", accuracy_parameter=1.0, group_width=width)
    if leaf_bits != 10:
        sim.set_tuning(gsim.TUNE_BH_LEAF_BITS, leaf_bits)
    if box:
        sim.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
    reference_tree = leaf_bits == 10 and not box
    sim.step(max(warmup, 1))
    sim.synchronize()
    sim.set_timing(True)
    sim.timing(reset=True)
    V = A = 0
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    V, A = (int(x) for x in sim.dump(gsim.DUMP_COUNTERS))  # last step's counts
    nodes = int(sim.dump(gsim.DUMP_NODE_TOTAL)[0])
    t = sim.timing(reset=True)
    sim.close()
    ms_step = elapsed * 1e3 / steps
    trav_ms = t["force_ms"] / max(t["steps"], 1)
    build_ms = t["build_ms"] / max(t["steps"], 1)
    flops = BH_FLOP_VISIT * V + BH_FLOP_ACCEPT * A
    achieved = flops / (trav_ms * 1e-3) / 1e12
    lane_ops = BH_LANEOPS_VISIT * V + BH_LANEOPS_ACCEPT * A
    # step roofline: build bytes at HBM peak + traversal lane-ops at VALU peak.
    # Build bytes 64 per particle (keys 12, one sort pass 16, leaf emission 36)
    # + 20 per cell node (nodes - particles: none leave the domain here)
    na = (BH_N + 63) // 64 * 64
    build_bytes = 64 * na + 20 * max(nodes - na, 0)
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
        "visits": V,
        "accepts": A,
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
            "valu_lane_op_frac": lane_ops / (trav_ms * 1e-3) / VALU_LANE_OPS,
            "build_bytes": build_bytes,
            "build_hbm_frac": build_bytes / (build_ms * 1e-3) / (PEAK_HBM_GBS * 1e9),
            "step_frac": t_star / (ms_step * 1e-3),
        },
    }


def batched_systems(count: int = 256, seed: int = 4):
    """The batched workload: `count` systems, N log-uniform in [1k, 64k],
    random-disk or exponential-disk (radius 100) chosen per system, masses
    U[1,10]; N < 8192 -> direct sum, else Barnes-Hut with theta = 0.7."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(64000))))
        kind = "random" if rng.uniform() < 0.5 else "exponential-disk"
        p = gsim.generate(kind, n, 100.0, 1.0, 10.0, 1.0, seed=1000 + k)
        algo = "direct-sum" if n < 8192 else "barnes-hut"
        out.append((p, algo, {"accuracy_parameter": 0.7} if algo == "barnes-hut" else {}))
    return out


def bench_batched(steps: int, serial_steps: int = 5) -> dict:
    systems = batched_systems()
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    batch = gsim.Batch(cap, max_systems=len(systems))
    for p, algo, kw in systems:
        batch.add(p, algo, **kw)
    waves = batch.autotune(steps_per_trial=2)
    batch.step(2)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.step(steps)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
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
    batched = len(systems) * steps / t_batch
    serial = len(systems) * serial_steps / t_serial
    return {
        "metric": "batched system-steps/s (256 mixed systems)",
        "value": batched,
        "unit": "system-steps/s",
        "higher_is_better": True,
        "ms_per_step": t_batch * 1e3 / steps,
        "one_by_one_value": serial,
        "speedup_vs_one_by_one": batched / serial,
        "particles": int(sum(p.n for p, _, _ in systems)),
        "config": {"workload": f"{len(systems)} systems, N log-uniform [1k, 64k], random or "
                               f"exponential disk (radius 100), N < 8192 direct sum ({n_ds}), "
                               f"else Barnes-Hut theta=0.7 ({len(systems) - n_ds}); one hipGraph "
                               f"replay per step", "steps": steps,
                   "one_by_one_steps": serial_steps, "waves_per_simd": waves},
    }


def served_requests(seconds: float, rate: float = 50.0, seed: int = 7):
    """The served workload: Poisson arrivals at `rate`/s for `seconds`; N
    log-uniform in [2k, 2M]; preset random, galaxy or galaxy-collision (size
    100, masses U[1,10]); Barnes-Hut with theta in {0.5, 0.7, 1.0}; steps
    uniform in {10..1000}.  Returns (arrival time, particles, theta, steps)."""
    rng = np.random.default_rng(seed)
    kinds = ("random", "galaxy", "galaxy-collision")
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        n = int(np.exp(rng.uniform(np.log(2000), np.log(2_000_000))))
        kind = kinds[int(rng.integers(3))]
        theta = float((0.5, 0.7, 1.0)[int(rng.integers(3))])
        steps = int(rng.integers(10, 1001))
        p = gsim.generate(kind, n, 100.0, 1.0, 10.0, 1.0, seed=50_000 + len(out))
        out.append((t, p, theta, steps))


def bench_serve(seconds: float, budget_gb: float = 200.0, max_active: int = 0,
                timeout_s: float = 900.0) -> dict:
    import threading

    reqs = served_requests(seconds)
    offered = sum(p.n * steps for _, p, _, steps in reqs)
    log(f"  {len(reqs)} requests, {offered / 1e9:.1f} G particle-steps offered over {seconds:g} s")
    srv = gsim.Server(int(budget_gb * 1e9), max_systems=2048, max_active_particles=max_active)
    rids = []

    def client():
        t0 = time.perf_counter()
        for at, p, theta, steps in reqs:
            wait = at - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            rids.append(srv.submit(p, steps, "barnes-hut", accuracy_parameter=theta))

    th = threading.Thread(target=client)
    t0 = time.perf_counter()
    th.start()
    while th.is_alive():  # arrivals: a progress line every 20 s of them
        th.join(20.0)
        if th.is_alive():
            st = srv.stats()
            log(f"  arrivals: {len(rids)} submitted, {st['completed']} done")
    # drain in slices, reporting what has completed, until idle or timeout_s
    deadline = time.perf_counter() + timeout_s
    drained = False
    while not drained and time.perf_counter() < deadline:
        drained = srv.drain(min(30.0, max(deadline - time.perf_counter(), 0.0)))
        if not drained:
            st = srv.stats()
            log(f"  draining: {st['completed']} of {len(rids)} done")
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat, queue = [], []
    for rid in rids:
        if srv.wait(rid, 0.0):
            _, info = srv.result(rid)
            lat.append(info["latency_ms"])
            queue.append(info["queue_ms"])
    srv.close()
    span = max(st["span_s"], 1e-9)
    inter = st["bh_interactions"] + st["ds_interactions"]
    return {
        "metric": "served systems/s (continuous batching)",
        "value": st["completed"] / span,
        "unit": "systems/s",
        "higher_is_better": True,
        "completed": st["completed"],
        "drained": drained,
        "p50_ms": float(np.percentile(lat, 50)) if lat else None,
        "p99_ms": float(np.percentile(lat, 99)) if lat else None,
        "queue_p50_ms": float(np.percentile(queue, 50)) if queue else None,
        "interactions_per_s": inter / span,
        "particle_steps_per_s": offered / span if drained else None,
        "offered_particle_steps_per_s": offered / seconds,
        "span_s": span,
        "wall_s": wall,
        "batch_steps": st["batch_steps"],
        "config": {"workload": f"Poisson arrivals 50/s for {seconds:g} s ({len(reqs)} requests), "
                               "N log-uniform [2k, 2M], preset random/galaxy/galaxy-collision "
                               "(size 100, masses U[1,10]), Barnes-Hut theta in {0.5,0.7,1.0}, "
                               f"steps U{{10..1000}}, HBM budget {budget_gb:g} GB",
                   "capacity_particles": st["capacity_particles"],
                   "max_active_particles": max_active},
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


def cpu_baseline(budget_s: float = 12.0) -> dict:
    """The oracle's FP32 direct sum (scalar C, one thread per core slice) on
    the same 50k workload, timed on this host: a bounded sample of whole
    force passes over all 50,000 x 50,000 pairs on up to 16 threads, and one
    pass on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline

    threads = max(1, min(16, os.cpu_count() or 1))
    p = gsim.generate("exponential-disk", DS_N, 100.0, 1.0, 10.0, 1.0, seed=2)
    eps2 = oracle.square_f32(0.2)
    # full passes over a target subset, sized to fit the time budget
    t0 = time.perf_counter()
    oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=threads)
    one = time.perf_counter() - t0
    reps = max(1, int(budget_s / max(one, 1e-3)))
    reps = min(reps, 8)
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=threads)
    el = time.perf_counter() - t0
    pairs = float(DS_N) * DS_N * reps
    t0 = time.perf_counter()
    oracle.ds_accel_f32(p.pos, p.mass, eps2, nthreads=1)
    el1 = time.perf_counter() - t0
    return {"value": pairs / el / 1e9, "unit": "Ginteractions/s", "cores": threads,
            "kind": "port",
            "one_core_value": float(DS_N) * DS_N / el1 / 1e9,
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"{reps + 1} full 50,000x50,000 FP32 force passes of the oracle "
                      f"({threads} threads) and one on 1 thread, integrate excluded"}


def cpu_baseline_bh(steps: int = 3, width: int = 32) -> dict:
    """The oracle's whole Barnes-Hut step (dense-pyramid build, group walk,
    integrate) on the same 500k workload and group width, on up to 16
    threads: ms per step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline

    threads = max(1, min(16, os.cpu_count() or 1))
    p = gsim.generate("two-exponential-disks", BH_N, 33.0, 1.0, 10.0, 1.0, seed=3).padded()
    eps2 = oracle.square_f32(0.2)
    theta2 = oracle.square_f32(1.0)
    pos, vel, mass = p.pos, p.vel, p.mass
    pos, vel, mass, _, _ = oracle.bh_step(pos, vel, mass, 1.0, 1e-3, eps2, theta2, width, threads)
    t0 = time.perf_counter()
    for _ in range(steps):
        pos, vel, mass, _, _ = oracle.bh_step(pos, vel, mass, 1.0, 1e-3, eps2, theta2, width,
                                              threads)
    el = time.perf_counter() - t0
    return {"value": el * 1e3 / steps, "unit": "ms/step", "cores": threads, "kind": "port",
            "sample": f"{steps + 1} whole Barnes-Hut steps of the oracle at 500,032 "
                      f"particles, W={width} ({threads} threads)"}


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
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival window of the served workload (0: skip)")
    ap.add_argument("--serve-max-active", type=int, default=0,
                    help="cap on particles stepping at once in the served run (0: none)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run known-size device copies (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the simulation step has no multi-GPU path; "
              "run with --gpus 1.", file=sys.stderr)
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
            f"{bh['traverse_ms']:.4f})")
    bt = None
    if not args.no_batch:
        log(f"batched 256 systems: steps {args.batch_steps}")
        bt = bench_batched(args.batch_steps)
        log(f"  {bt['value']:.0f} system-steps/s ({bt['ms_per_step']:.3f} ms/step), "
            f"{bt['speedup_vs_one_by_one']:.2f}x one by one")
    sv = None
    if args.serve_seconds > 0:
        log(f"served: {args.serve_seconds:g} s of arrivals")
        sv = bench_serve(args.serve_seconds, max_active=args.serve_max_active)
        log(f"  {sv['value']:.2f} systems/s, p50 {sv['p50_ms']} ms, p99 {sv['p99_ms']} ms, "
            f"{sv['interactions_per_s'] / 1e9:.1f} Ginteractions/s")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu = cpu_baseline()
        log(f"  {cpu['value']:.2f} Ginteractions/s on {cpu['cores']} threads, "
            f"{cpu['one_core_value']:.3f} on one")
        if bh is not None:
            bh["cpu_baseline"] = cpu_baseline_bh(width=args.bh_width)
            log(f"  Barnes-Hut 500k: {bh['cpu_baseline']['value']:.1f} ms/step on "
                f"{bh['cpu_baseline']['cores']} threads")
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
                   "particles": DS_N, "global_batch": 1, "parallelism": "single-gpu"},
        "pcie_inclusive_ms_one_step": ds["pcie_inclusive_ms"],
        "roofline": ds["roofline"],
        "cpu_baseline": cpu,
        "secondary": [x for x in (bh, bt, sv) if x],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
