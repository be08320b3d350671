"""The batched and served benchmark workloads, checked on the GPU.

Batched: the 256-system mix the bench times (N log-uniform in [1k, 64k];
direct sum below 8,192 particles, Barnes-Hut with theta = 0.7 above), stepped
together; every system must equal its alone run bit for bit, and sampled
systems of both algorithms are checked step by step against the oracle.

Served: the request stream the bench serves (Poisson 50/s, N log-uniform in
[2k, 2M], presets random / galaxy / galaxy-collision, theta in {0.5, 0.7,
1.0}, 10-1000 steps) for a one-second arrival window, plus a request with
G = 0.5 and one direct-sum request, admitted under a budget that makes late
arrivals wait for space; every result must equal its alone run bit for bit,
and sampled requests (theta = 0.5, G = 0.5, the direct sum) are checked
against the oracle on their first and last steps.
"""
import threading
import time

import pytest

import gsim
from gsim import workloads as WL
from gpu_checks import same, say, step_alone_checked

pytestmark = pytest.mark.gpu


def _pick_batched_samples(systems):
    """Four direct-sum and four Barnes-Hut systems: the smallest and largest
    of each algorithm and two from the middle."""
    out = []
    for algo in ("direct-sum", "barnes-hut"):
        idx = sorted((p.n, k) for k, (p, a, _) in enumerate(systems) if a == algo)
        m = len(idx)
        out += [idx[0][1], idx[m // 3][1], idx[2 * m // 3][1], idx[-1][1]]
    return out


def test_batched_256_system_workload():
    systems = WL.batched_systems()
    assert len(systems) == 256
    n_ds = sum(1 for _, a, _ in systems if a == "direct-sum")
    assert 0 < n_ds < 256
    steps = 2
    padded = [(p.padded(), a, kw) for p, a, kw in systems]
    cap = sum(p.n for p, _, _ in padded)
    say(f"batched: 256 systems ({n_ds} direct sum), {cap} padded particles, {steps} steps")
    with gsim.Batch(cap, max_systems=len(systems)) as b:
        ids = [b.add(p, a, **kw) for p, a, kw in systems]
        b.step(steps)
        got = [b.read(sid) for sid in ids]
    samples = set(_pick_batched_samples(systems))
    for k, ((p, algo, kw), (pp, _, _), g) in enumerate(zip(systems, padded, got)):
        if k in samples:
            say(f"  system {k}: {algo} n={p.n} {kw}, checked against the oracle")
            alone, recs = step_alone_checked(pp, algo, kw, steps, check=(0, 1))
            assert len(recs) == 2
        else:
            with gsim.Simulation(pp, algo, **kw) as s:
                s.step(steps)
                alone = s.read()
        same(g, gsim.Particles(alone.pos[: p.n], alone.vel[: p.n], alone.mass[: p.n]))


def _served_mix():
    """(arrival s, particles, algorithm, parameters, steps) per request."""
    stream = WL.served_requests(1.0)
    reqs = [(r.at, r.particles, "barnes-hut", {"accuracy_parameter": r.theta}, r.steps)
            for r in stream]
    end = stream[-1].at
    extra_bh = gsim.generate("galaxy", 40_000, 100.0, 1.0, 10.0, 0.5, seed=61)
    reqs.append((end, extra_bh, "barnes-hut", {"accuracy_parameter": 0.7,
                                               "gravitational_const": 0.5}, 37))
    extra_ds = gsim.generate("galaxy-collision", 6_000, 100.0, 1.0, 10.0, 1.0, seed=62)
    reqs.append((end, extra_ds, "direct-sum", {"softening_len": 0.3}, 25))
    return reqs


def test_served_workload_window():
    mix = _served_mix()
    arrivals = [m[0] for m in mix]
    reqs = [m[1:] for m in mix]
    thetas = {kw.get("accuracy_parameter") for _, a, kw, _ in reqs if a == "barnes-hut"}
    assert {0.5, 0.7, 1.0} <= thetas
    largest = max(p.n for p, _, _, _ in reqs)
    total = sum(p.n for p, _, _, _ in reqs)
    # room for the largest request and about a third of the rest at once,
    # so later arrivals queue until earlier ones retire
    budget_particles = max(largest + 64, total // 3)
    say(f"served: {len(reqs)} requests, largest {largest}, total {total} particles, "
        f"budget {budget_particles}")
    budget = budget_particles * gsim.Batch.bytes_per_particle()
    results = [None] * len(reqs)
    with gsim.Server(budget, max_systems=256) as srv:
        rids = [None] * len(reqs)

        def client():
            t0 = time.perf_counter()
            for k, (p, algo, kw, steps) in enumerate(reqs):
                wait = arrivals[k] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                rids[k] = srv.submit(p, steps, algo, **kw)

        th = threading.Thread(target=client)
        th.start()
        th.join()
        assert srv.drain(300.0), "server did not drain"
        for k, rid in enumerate(rids):
            assert srv.wait(rid, 0.0)
            results[k], info = srv.result(rid)
            assert info["steps"] == reqs[k][3]
        st = srv.stats()
        say(f"  served {st['completed']} in {st['span_s']:.2f} s, p50 {st['p50_ms']:.0f} ms, "
            f"p99 {st['p99_ms']:.0f} ms")
        assert st["completed"] == len(reqs) and st["failed"] == 0
    # sampled requests against the oracle: the first theta = 0.5 request
    # below 300k particles, the G = 0.5 request and the direct-sum request
    half = [k for k, (p, a, kw, _) in enumerate(reqs)
            if kw.get("accuracy_parameter") == 0.5 and p.n < 300_000]
    assert half, "the window holds no theta = 0.5 request below 300k particles"
    checked = {half[0], len(reqs) - 2, len(reqs) - 1}
    for k, (p, algo, kw, steps) in enumerate(reqs):
        pp = p.padded()
        if k in checked:
            say(f"  request {k}: {algo} n={p.n} {kw} steps={steps}, oracle on first and last")
            alone, _ = step_alone_checked(pp, algo, kw, steps, check=(0, steps - 1))
        else:
            with gsim.Simulation(pp, algo, **kw) as s:
                s.step(steps)
                alone = s.read()
        same(results[k], gsim.Particles(alone.pos[: p.n], alone.vel[: p.n], alone.mass[: p.n]))
