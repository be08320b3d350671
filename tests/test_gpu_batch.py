"""Batched form: many systems stepped together through gsim_batch_* must give
each system exactly (bit for bit) what stepping it alone gives."""
import sys

import numpy as np
import pytest

import gsim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _systems():
    """A mixed set: small direct-sum systems, Barnes-Hut systems of uneven
    size and parameters, one with particles leaving the domain."""
    out = []
    rng = np.random.default_rng(11)
    specs = [("random", 1000, "direct-sum", {}), ("galaxy", 3000, "direct-sum", {"softening_len": 0.5}),
             ("exponential-disk", 7000, "direct-sum", {}),
             ("galaxy-collision", 9000, "barnes-hut", {"accuracy_parameter": 0.7}),
             ("exponential-disk", 20000, "barnes-hut", {"accuracy_parameter": 0.7}),
             ("random", 12345, "barnes-hut", {"accuracy_parameter": 1.0, "simulation_speed": 2.0}),
             ("galaxy", 40000, "barnes-hut", {"accuracy_parameter": 0.5, "gravitational_const": 0.5}),
             ("galaxy-collision", 15000, "barnes-hut", {"accuracy_parameter": 0.7, "group_width": 32}),
             ("exponential-disk", 4321, "barnes-hut", {"group_width": 32, "simulation_speed": 2.0})]
    for k, (kind, n, algo, kw) in enumerate(specs):
        p = gsim.generate(kind, n, 100.0, 1.0, 10.0, kw.get("gravitational_const", 1.0), seed=100 + k)
        if k == 5:
            p.pos[:7] = rng.uniform(600, 900, size=(7, 2))  # killed in the first build
        out.append((p, algo, kw))
    return out


def _alone(p, algo, kw, steps):
    with gsim.Simulation(p, algo, **kw) as s:
        s.step(steps)
        return s.read()


def _same(a, b):
    assert np.array_equal(a.pos, b.pos)
    assert np.array_equal(a.vel, b.vel)
    assert np.array_equal(a.mass, b.mass)


def test_batch_matches_single_system_runs():
    systems = _systems()
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap + 4096, max_systems=16) as b:
        ids = [b.add(p, algo, **kw) for p, algo, kw in systems]
        b.step(6)
        for sid, (p, algo, kw) in zip(ids, systems):
            _say(f"batch vs alone: {algo} n={p.n} {kw}")
            _same(b.read(sid), _alone(p, algo, kw, 6))


def test_batch_many_sort_tiles_matches_single_system_runs():
    """4.5 million Barnes-Hut positions: 2,198 radix tiles, more than the GPU
    holds at once, taken by persistent workgroups over several rounds."""
    systems = [(gsim.generate("two-exponential-disks", 1500000, 33.0, 1.0, 10.0, seed=31 + k),
                "barnes-hut", {"accuracy_parameter": 0.7 + 0.1 * k}) for k in range(3)]
    systems.append((gsim.generate("random", 2000, 50.0, 1.0, 10.0, seed=34), "direct-sum", {}))
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap + 4096, max_systems=8) as b:
        ids = [b.add(p, algo, **kw) for p, algo, kw in systems]
        b.step(2)
        for sid, (p, algo, kw) in zip(ids, systems):
            _say(f"batch vs alone: {algo} n={p.n} {kw}")
            _same(b.read(sid), _alone(p, algo, kw, 2))


def test_batch_membership_changes_between_steps():
    systems = _systems()
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap, max_systems=8) as b:
        a = b.add(*systems[0][:2], **systems[0][2])
        c = b.add(*systems[4][:2], **systems[4][2])
        d = b.add(*systems[3][:2], **systems[3][2])
        b.step(3)
        b.remove(d)
        e = b.add(*systems[6][:2], **systems[6][2])   # joins mid-run
        f = b.add(*systems[1][:2], **systems[1][2])
        b.step(4)
        _same(b.read(a), _alone(*systems[0][:2], systems[0][2], 7))
        _same(b.read(c), _alone(*systems[4][:2], systems[4][2], 7))
        _same(b.read(e), _alone(*systems[6][:2], systems[6][2], 4))
        _same(b.read(f), _alone(*systems[1][:2], systems[1][2], 4))
        info = b.info()
        assert info["systems"] == 4 and info["steps"] == 7


def test_batch_autotune_leaves_state_unchanged():
    systems = _systems()[2:6]
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap, max_systems=8) as b:
        ids = [b.add(p, algo, **kw) for p, algo, kw in systems]
        b.step(2)
        w = b.autotune(steps_per_trial=2)
        t = b.tuning()
        assert 1 <= w <= 8 and t["waves_per_simd"] == w
        assert t["sort_tile"] in (2048, 3072, 4096) and 1 <= t["run_queue"] <= 64
        b.step(3)
        for sid, (p, algo, kw) in zip(ids, systems):
            _same(b.read(sid), _alone(p, algo, kw, 5))


@pytest.mark.parametrize("knobs", [
    {"sort_tile": 2048}, {"sort_tile": 4096}, {"run_queue": 1}, {"run_queue": 7},
    {"prio_step": 0, "prio_carry": 0}, {"prio_step": 5, "prio_carry": 3},
    {"waves_per_simd": 1, "sort_tile": 2048, "run_queue": 16}])
def test_batch_tuning_knobs_change_no_result(knobs):
    """Every launch knob the autotuner turns leaves each system bit-identical
    to its alone run (sort tiles, run-queue cuts and issue priority change
    only how the work is scheduled), including tiles that split systems and
    queues cut after a single run."""
    systems = _systems()
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap, max_systems=16) as b:
        ids = [b.add(p, algo, **kw) for p, algo, kw in systems]
        b.set_tuning(**knobs)
        assert all(b.tuning()[k] == v for k, v in knobs.items())
        b.step(3)
        for sid, (p, algo, kw) in zip(ids, systems):
            _say(f"tuning {knobs}: {algo} n={p.n} {kw}")
            _same(b.read(sid), _alone(p, algo, kw, 3))


def test_batch_tuning_rejects_bad_values():
    with gsim.Batch(8192, max_systems=2) as b:
        for bad in ({"sort_tile": 1000}, {"run_queue": 0}, {"run_queue": 65},
                    {"waves_per_simd": 9}, {"prio_step": 5000}):
            with pytest.raises(gsim.GsimError) as e:
                b.set_tuning(**bad)
            assert e.value.status == 1


def test_batch_capacity_is_enforced():
    p = gsim.generate("random", 5000, 50.0, 1.0, 2.0, seed=1)
    with gsim.Batch(8192, max_systems=2) as b:
        b.add(p, "barnes-hut")
        with pytest.raises(gsim.GsimError) as e:
            b.add(p, "barnes-hut", group_width=48)  # votes are 32 or 64 wide
        assert e.value.status == 1
        with pytest.raises(gsim.GsimError) as e:
            b.add(p, "direct-sum")  # 5056 more padded particles do not fit in 8192
        assert e.value.status == 6
    with gsim.Batch(40000, max_systems=4) as b:
        big = gsim.generate("random", 16321, 50.0, 1.0, 2.0, seed=3)  # 16,384 padded
        with pytest.raises(gsim.GsimError) as e:
            b.add(big, "direct-sum")  # stepped alone with the symmetric kernel
        assert e.value.status == 1
        b.add(big, "barnes-hut")
    with gsim.Batch(8192, max_systems=2) as b:
        b.add(p, "barnes-hut")
        small = gsim.generate("random", 100, 50.0, 1.0, 2.0, seed=2)
        b.add(small, "direct-sum")
        with pytest.raises(gsim.GsimError):
            b.add(small, "direct-sum")  # max_systems


def _escaping(n=2000, seed=5):
    """Barnes-Hut input whose particle 0 crosses the domain edge on step 1
    (x = 499.5, vx = 1000, dt = 1e-3), so step 2 kills it."""
    p = gsim.generate("random", n, 100.0, 1.0, 10.0, seed=seed)
    p.pos[0] = (499.5, 0.0)
    p.vel[0] = (1000.0, 0.0)
    return p


def test_batch_step_budgets_freeze_exactly():
    systems = _systems()[:6] + [(_escaping(), "barnes-hut", {}), (_escaping(seed=6), "barnes-hut", {})]
    budgets = [1, 3, None, 2, 5, 4, 1, 2]
    cap = sum((p.n + 63) // 64 * 64 for p, _, _ in systems)
    with gsim.Batch(cap, max_systems=16) as b:
        ids = []
        for (p, algo, kw), k in zip(systems, budgets):
            sid = b.add(p, algo, **kw)
            b.set_steps(sid, k)
            ids.append(sid)
        b.step(3)
        b.step(4)
        for sid, (p, algo, kw), k in zip(ids, systems, budgets):
            want = 7 if k is None else min(k, 7)
            _say(f"budget {k}: {algo} n={p.n} {kw}")
            assert b.steps_done(sid) == want
            _same(b.read(sid), _alone(p, algo, kw, want))
        # extend a spent budget: the system resumes from where it stopped
        b.set_steps(ids[0], 2)
        b.step(5)
        assert b.steps_done(ids[0]) == 3
        p, algo, kw = systems[0]
        _same(b.read(ids[0]), _alone(p, algo, kw, 3))
    esc = _alone(_escaping(), "barnes-hut", {}, 1)
    assert esc.pos[0, 0] > 500.0 and esc.mass[0] > 0.0  # left the domain, not yet killed


def test_batch_accepts_match_single_path_counters():
    p = gsim.generate("galaxy", 20000, 100.0, 1.0, 10.0, seed=21)
    with gsim.Simulation(p, "barnes-hut", accuracy_parameter=0.7) as s:
        want_v = want_a = 0
        for _ in range(3):
            s.step(1)
            v, a = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
            want_v += v
            want_a += a
    with gsim.Batch(32768, max_systems=4) as b:
        sid = b.add(p, "barnes-hut", accuracy_parameter=0.7)
        b.set_steps(sid, 3)
        b.step(5)  # the two steps past the budget add nothing
        assert b.accepts() == want_a
        assert b.visits() == want_v
    # autotuner trials run on a snapshot: they leave both totals untouched
    with gsim.Batch(32768, max_systems=4) as b:
        sid = b.add(p, "barnes-hut", accuracy_parameter=0.7)
        b.step(3)
        b.autotune(steps_per_trial=2)
        assert (b.visits(), b.accepts()) == (want_v, want_a)
