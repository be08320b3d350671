"""Serving: requests admitted into one continuously-stepped batch must each
get exactly (bit for bit) what running the system alone gives, whatever else
shares the batch and whenever it arrives."""
import sys
import threading
import time

import numpy as np
import pytest

import gsim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _alone(p, algo, kw, steps):
    with gsim.Simulation(p, algo, **kw) as s:
        s.step(steps)
        return s.read()


def _same(a, b):
    assert np.array_equal(a.pos, b.pos)
    assert np.array_equal(a.vel, b.vel)
    assert np.array_equal(a.mass, b.mass)


def _requests():
    specs = [("random", 700, "direct-sum", {}, 5),
             ("galaxy", 5000, "barnes-hut", {"accuracy_parameter": 0.7}, 12),
             ("galaxy-collision", 3000, "direct-sum", {"softening_len": 0.5}, 9),
             ("exponential-disk", 30000, "barnes-hut", {"accuracy_parameter": 0.5}, 20),
             ("random", 12000, "barnes-hut", {"accuracy_parameter": 1.0, "simulation_speed": 2.0}, 3),
             ("two-exponential-disks", 9000, "barnes-hut", {}, 17),
             ("galaxy", 2000, "direct-sum", {}, 1),
             ("random", 40000, "barnes-hut", {"accuracy_parameter": 0.7}, 8),
             ("galaxy", 8000, "barnes-hut", {"group_width": 32}, 11)]
    out = []
    for k, (kind, n, algo, kw, steps) in enumerate(specs):
        p = gsim.generate(kind, n, 100.0, 1.0, 10.0, seed=300 + k)
        out.append((p, algo, kw, steps))
    return out


def _serve(reqs, budget_particles, max_active=0, stagger_s=0.0, arena=0):
    """Submits every request (from a second thread, `stagger_s` apart) and
    returns results in submission order."""
    budget = budget_particles * gsim.Batch.bytes_per_particle()
    with gsim.Server(budget, max_systems=64, max_active_particles=max_active,
                     arena_particles=arena) as srv:
        rids = [None] * len(reqs)

        def client():
            for k, (p, algo, kw, steps) in enumerate(reqs):
                rids[k] = srv.submit(p, steps, algo, **kw)
                if stagger_s:
                    time.sleep(stagger_s)

        th = threading.Thread(target=client)
        th.start()
        th.join()
        assert srv.drain(120.0)
        out = []
        for rid in rids:
            assert srv.wait(rid, 10.0)
            out.append(srv.result(rid))
        stats = srv.stats()
    return out, stats


def test_server_results_match_alone_runs():
    reqs = _requests()
    want = [_alone(p, algo, kw, steps) for p, algo, kw, steps in reqs]
    total = sum((p.n + 63) // 64 * 64 for p, _, _, _ in reqs)
    # roomy budget, staggered arrivals: requests join a batch already running
    _say("serve: staggered arrivals, everything fits")
    got, stats = _serve(reqs, total + 4096, stagger_s=0.003)
    for (res, info), w, (p, algo, kw, steps) in zip(got, want, reqs):
        _say(f"  {algo} n={p.n} steps={steps} latency={info['latency_ms']:.2f} ms")
        assert info["n"] == p.n and info["steps"] == steps
        _same(res, w)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    assert 0.0 < stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    assert stats["bh_interactions"] > 0 and stats["ds_interactions"] > 0
    # tight budget: only a few requests fit at once, the rest queue
    _say("serve: tight budget, requests queue for space")
    got, stats = _serve(reqs, 48 * 1024)
    for (res, info), w in zip(got, want):
        _same(res, w)
    assert stats["completed"] == len(reqs)
    # active-particle cap of one system at a time
    _say("serve: active-particle cap")
    got, _ = _serve(reqs[:4], total, max_active=1)
    for (res, info), w in zip(got, want[:4]):
        _same(res, w)


def test_server_spreads_requests_over_arenas():
    """A budget split into three arenas (each with its own stream, stepped
    together): requests land in whichever arena has room and every result is
    still its alone run's, bit for bit; a request larger than the largest
    arena is refused at submission."""
    reqs = _requests()
    want = [_alone(p, algo, kw, steps) for p, algo, kw, steps in reqs]
    arena = 40960  # the largest request, 40,000 particles, fits one arena
    _say("serve: three arenas")
    got, stats = _serve(reqs, 3 * arena, arena=arena, stagger_s=0.002)
    for (res, info), w, (p, algo, kw, steps) in zip(got, want, reqs):
        _same(res, w)
    assert stats["arenas"] == 3 and stats["capacity_particles"] == 3 * arena
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    big = gsim.generate("random", arena + 1, 100.0, 1.0, 10.0, seed=12)
    budget = 3 * arena * gsim.Batch.bytes_per_particle()
    with gsim.Server(budget, max_systems=8, arena_particles=arena) as srv:
        with pytest.raises(gsim.GsimError) as e:
            srv.submit(big, 2, "barnes-hut")
        assert e.value.status == 6


def test_server_edge_cases():
    p = gsim.generate("random", 1000, 100.0, 1.0, 10.0, seed=9)
    with gsim.Server(8192 * gsim.Batch.bytes_per_particle(), max_systems=4) as srv:
        big = gsim.generate("random", 9000, 100.0, 1.0, 10.0, seed=10)
        with pytest.raises(gsim.GsimError) as e:
            srv.submit(big, 3, "barnes-hut")  # more than the whole budget
        assert e.value.status == 6
        rid = srv.submit(p, 0, "direct-sum")  # no steps: the input comes back
        assert srv.wait(rid, 5.0)
        res, info = srv.result(rid)
        _same(res, p)
        rid = srv.submit(p, 4, "barnes-hut")
        with pytest.raises(gsim.GsimError):
            srv.result(rid + 1000)  # unknown id
        assert srv.wait(rid, 30.0)
        res, _ = srv.result(rid)
        _same(res, _alone(p, "barnes-hut", {}, 4))
