"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Tolerances (FP32 forces): the GPU's v_rsq_f32 and FMA differ from the
oracle's IEEE 1/sqrt and unfused sums by a few ulp per interaction, so forces
are compared with
    system-relative  ||a_gpu - a_ref|| / ||a_ref||            <= 1e-5
    per particle     |a_gpu,i - a_ref,i| / max(|a_ref,i|, rms) <= 1e-5
and where the GPU adds the same terms in another order (the direct sums, the
Barnes-Hut walk's four partials) the per-particle bound is against the same
terms summed in double: no worse than 1.5x the FP32 oracle's own error.
Integer work (cells, permutation, tree topology, visit/accept counts) and the
centre-of-mass arithmetic are compared bit for bit.
"""
import sys

import numpy as np
import pytest

import gsim
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
EPS2 = H.sq32(0.2)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(p, algo, **kw):
    s = gsim.Simulation(p, algo, **kw)
    s.set_debug(gsim.DEBUG_ACCEL)
    return s


def _force_vs_f64(a, ref32, ref64, tol=FORCE_TOL):
    """Different summation order than the reference's sequential j loop: the
    system error against the FP32 oracle stays <= tol, and per particle the
    GPU is no less accurate (against FP64) than the FP32 oracle itself, whose
    own 50k-term sums deviate by ~1e-5 at worst."""
    sysr = H.system_rel_err(a, ref32)
    assert sysr <= tol, f"system rel err {sysr:.3e}"
    eg = H.per_particle_rel(a, ref64)
    e32 = H.per_particle_rel(ref32, ref64)
    assert eg.max() <= max(1.5 * e32.max(), 1e-6), (eg.max(), e32.max())
    assert np.quantile(eg, .999) <= max(1.5 * np.quantile(e32, .999), 1e-6)
    return sysr, eg.max(), e32.max()


def _bh_force_ok(a, ref32, ref64, tol=FORCE_TOL):
    """A walk adds the same accepted terms as the oracle's, in its own order
    (64-node run chunks into four partials): against the FP32 oracle the
    system error stays <= tol, and per particle the GPU is no less accurate
    (against the same terms summed in double) than the oracle's own FP32
    sequential sum."""
    return _force_vs_f64(a, ref32, ref64, tol)


def _force_ok(a, ref, tol=FORCE_TOL):
    sysr = H.system_rel_err(a, ref)
    per = H.per_particle_rel(a, ref)
    assert sysr <= tol, f"system rel err {sysr:.3e}"
    assert per.max() <= tol, f"per-particle max {per.max():.3e}, p99.9 {np.quantile(per, .999):.3e}"
    return sysr, per.max()


# ---- direct sum --------------------------------------------------------------------

def test_ds_config1_4096_100_steps():
    """4096 particles, uniform disk r=100, v=0, m in [1,10], G=1, eps=0.2,
    dt=1e-3, 100 steps: one-step forces vs the FP32 oracle, and the 100-step
    trajectory vs FP64 no worse than the FP32 oracle's own deviation x 4."""
    _say("ds config1")
    p = gsim.generate("random", 4096, 100.0, 1.0, 10.0, seed=101)
    s = _sim(p, "direct-sum")
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)[: p.n]
    ref = O.ds_accel_f32(p.pos, p.mass, EPS2)
    _force_ok(a, ref)
    a64 = O.ds_accel_f64(p.pos, p.mass, EPS2)
    assert H.system_rel_err(a, a64) < 2e-6
    s.step(99)
    g = s.read()
    s.close()
    po, vo = p.pos.copy(), p.vel.copy()
    for _ in range(100):
        po, vo = O.ds_step(po, vo, p.mass, 1.0, 1e-3, EPS2)
    p64, v64 = O.ds_steps_f64(p.pos, p.vel, p.mass, 1.0, 1e-3, EPS2, 100)
    dev32 = np.linalg.norm(po - p64)
    devg = np.linalg.norm(g.pos - p64)
    _say(f"  trajectory deviation from FP64: oracle {dev32:.3e}, gpu {devg:.3e}")
    assert devg <= 4.0 * dev32 + 1e-6
    assert np.linalg.norm(g.vel - v64) <= 4.0 * np.linalg.norm(vo - v64) + 1e-6
    assert np.array_equal(g.mass, p.mass)


@pytest.mark.parametrize("R,S,rcp", [(1, 0, 0), (1, 1, 0), (2, 8, 0), (4, 3, 0), (2, 32, 1),
                                     (1, 48, 1)])
def test_ds_50k_one_step_all_shapes(R, S, rcp):
    """50,000-particle exponential disk (the direct-sum benchmark workload):
    forces of every launch shape within the tolerance of the FP32 oracle."""
    _say(f"ds 50k R={R} S={S} rcp={rcp}")
    p = gsim.generate("exponential-disk", 50000, 100.0, 1.0, 10.0, seed=102)
    s = _sim(p, "direct-sum")
    s.set_ds_shape(R, S)
    s.set_tuning(gsim.TUNE_DS_RCP, rcp)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)[: p.n]
    g = s.read()
    s.close()
    pp = p.padded()
    ref = O.ds_accel_f32(pp.pos, pp.mass, EPS2)[: p.n]
    ref64 = O.ds_accel_f64(pp.pos, pp.mass, EPS2)[: p.n]
    sysr, eg, e32 = _force_vs_f64(a, ref, ref64)
    _say(f"  system {sysr:.2e}; per-particle max vs FP64: gpu {eg:.2e}, oracle {e32:.2e}")
    # integrate: v' = v + a*(dt*G), p' = p + ((v+v')*0.5)*dt on the GPU's forces
    kv = np.float32(1e-3) * np.float32(1.0)
    nv = (p.vel + a * kv).astype(np.float32)
    assert np.array_equal(g.vel, nv)
    np_ = (p.pos + ((p.vel + nv) * np.float32(0.5)) * np.float32(1e-3)).astype(np.float32)
    assert np.array_equal(g.pos, np_)


@pytest.mark.parametrize("cols", [0, 1, 2])
def test_ds_50k_symmetric_every_tile_width(cols):
    """The symmetric kernel (the default at 50,000): each pair evaluated once,
    forces within the tolerance of the FP32 oracle for every tile width, and
    the integration applied to exactly those forces."""
    _say(f"ds 50k symmetric cols={cols}")
    p = gsim.generate("exponential-disk", 50000, 100.0, 1.0, 10.0, seed=102)
    s = _sim(p, "direct-sum")
    s.set_tuning(gsim.TUNE_DS_SYM_COLS, cols)
    s.set_tuning(gsim.TUNE_DS_SYMMETRIC, 1)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)[: p.n]
    g = s.read()
    s.close()
    pp = p.padded()
    ref = O.ds_accel_f32(pp.pos, pp.mass, EPS2)[: p.n]
    ref64 = O.ds_accel_f64(pp.pos, pp.mass, EPS2)[: p.n]
    sysr, eg, e32 = _force_vs_f64(a, ref, ref64)
    _say(f"  system {sysr:.2e}; per-particle max vs FP64: gpu {eg:.2e}, oracle {e32:.2e}")
    kv = np.float32(1e-3) * np.float32(1.0)
    nv = (p.vel + a * kv).astype(np.float32)
    assert np.array_equal(g.vel, nv)
    np_ = (p.pos + ((p.vel + nv) * np.float32(0.5)) * np.float32(1e-3)).astype(np.float32)
    assert np.array_equal(g.pos, np_)


def test_ds_symmetric_is_default_in_its_range_and_deterministic():
    p = gsim.generate("galaxy", 20000, 100.0, 1.0, 10.0, seed=105)
    runs = []
    for sym in (None, 1, 0):
        s = _sim(p, "direct-sum")
        if sym is not None:
            s.set_tuning(gsim.TUNE_DS_SYMMETRIC, sym)
        s.step(2)
        runs.append(s.read())
        s.close()
    assert np.array_equal(runs[0].vel, runs[1].vel)       # default = symmetric
    assert not np.array_equal(runs[0].vel, runs[2].vel)   # a different summation order
    assert np.allclose(runs[0].vel, runs[2].vel, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 500, 1000, 4103, 9000, 16384])
def test_ds_symmetric_ragged_sizes(n):
    """Block counts below, at and off a multiple of the row group (16 blocks),
    forced onto the symmetric kernel: forces against the oracle, then three
    steps against the oracle's trajectory."""
    _say(f"ds symmetric n={n}")
    p = gsim.generate("galaxy", n, 50.0, 1.0, 10.0, seed=n + 7)
    s = _sim(p, "direct-sum")
    s.set_tuning(gsim.TUNE_DS_SYMMETRIC, 1)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)[: p.n]
    pp = p.padded()
    ref = O.ds_accel_f32(pp.pos, pp.mass, EPS2)[: p.n]
    ref64 = O.ds_accel_f64(pp.pos, pp.mass, EPS2)[: p.n]
    if n > 1:
        _force_vs_f64(a, ref, ref64)
    else:
        assert np.array_equal(a, np.zeros_like(a))  # alone: only itself and massless pads
    s.step(2)
    g = s.read()
    s.close()
    po, vo = pp.pos, pp.vel
    for _ in range(3):
        po, vo = O.ds_step(po, vo, pp.mass, 1.0, 1e-3, EPS2)
    assert np.allclose(g.pos, po[:n], rtol=1e-6, atol=1e-6)
    assert np.allclose(g.vel, vo[:n], rtol=1e-4, atol=1e-4)


def test_ds_symmetric_zero_softening_is_nan():
    p = gsim.generate("random", 700, 10.0, 1.0, 2.0, seed=4)
    z = _sim(p, "direct-sum", softening_len=0.0)
    z.set_tuning(gsim.TUNE_DS_SYMMETRIC, 1)
    z.step(1)
    assert np.isnan(z.read().vel).all()  # the self term, as in the reference sum
    z.close()


def test_ds_ragged_and_tiny_sizes():
    for n in (1, 2, 63, 64, 65, 1000):
        _say(f"ds n={n}")
        p = gsim.generate("galaxy", n, 50.0, 1.0, 10.0, seed=n)
        s = _sim(p, "direct-sum")
        assert s.na == (n + 63) // 64 * 64
        s.step(3)
        g = s.read()
        s.close()
        pp = p.padded()
        po, vo = pp.pos, pp.vel
        for _ in range(3):
            po, vo = O.ds_step(po, vo, pp.mass, 1.0, 1e-3, EPS2)
        assert np.allclose(g.pos, po[:n], rtol=1e-6, atol=1e-6)
        assert np.allclose(g.vel, vo[:n], rtol=1e-4, atol=1e-4)


def test_ds_ignores_simulation_speed_and_zero_softening_is_nan():
    p = gsim.generate("random", 128, 10.0, 1.0, 2.0, seed=3)
    a = _sim(p, "direct-sum", simulation_speed=5.0)
    b = _sim(p, "direct-sum", simulation_speed=1.0)
    a.step(2); b.step(2)
    assert np.array_equal(a.read().pos, b.read().pos)  # SimShader.comp:68-69 use dt only
    a.close(); b.close()
    z = _sim(p, "direct-sum", softening_len=0.0)
    z.step(1)
    assert np.isnan(z.read().vel).all()
    z.close()


# ---- Barnes-Hut: build -----------------------------------------------------------------

def _bh_inputs():
    yield "dyadic-700", *H.dyadic_particles(700, seed=201)[::2]
    rng = np.random.default_rng(202)
    pos = rng.normal(0, 40, size=(3000, 2)).astype(np.float32)
    pos[:20] = rng.uniform(-700, 700, size=(20, 2))  # some leave the domain
    pos[20] = [np.nan, 0]
    yield "normal-3000-kills", pos, rng.uniform(1, 10, 3000).astype(np.float32)
    p = gsim.generate("galaxy-collision", 10000, 300.0, 1.0, 10.0, seed=203)
    yield "collision-10000", p.pos, p.mass
    p = gsim.generate("random", 5000, 0.5, 1.0, 10.0, seed=204)  # crowded leaves
    yield "crowded-5000", p.pos, p.mass
    p = gsim.generate("two-exponential-disks", 500000, 33.0, 1.0, 10.0, seed=205)
    yield "two-disks-500000", p.pos, p.mass


@pytest.mark.parametrize("case", list(range(5)))
def test_bh_tree_bit_exact(case):
    name, pos, mass = list(_bh_inputs())[case]
    _say(f"bh tree {name}")
    n = len(mass)
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = _sim(p, "barnes-hut")
    s.step(1)
    t = s.tree()
    cells = s.dump(gsim.DUMP_CELLS)
    m30 = s.dump(gsim.DUMP_MORTON30)
    s.close()
    pp = p.padded()
    ot, nn, nin, pk, mk = O.bh_build(pp.pos, pp.mass)
    assert t["total"] == nn
    assert np.array_equal(cells, H.cells_numpy(pp.pos))
    ok = cells != 0xFFFFFFFF
    assert np.array_equal(m30[ok] >> np.uint32(10), cells[ok])
    assert np.array_equal(t["sorted_src"], ot["sorted_src"])
    assert np.array_equal(t["counts"], ot["counts"])
    assert np.array_equal(t["radii"], ot["radii"])
    assert np.array_equal(t["node_mass"], ot["node_mass"])
    assert np.array_equal(t["node_pos"], ot["node_pos"])


@pytest.mark.parametrize("bits", [10, 15])
def test_bh_tree_bit_exact_many_sort_tiles(bits):
    """3,000,000 particles: 733 radix tiles of 4,096 keys (the tile size of
    builds above 786,432 particles), more than the GPU holds at once (two per
    CU), so tiles wait on predecessors that started in an earlier round of
    workgroups.  Every node array equals the oracle's sparse build
    (itself pinned to the dense restatement at 10 bits)."""
    _say(f"bh tree two-disks-3000000 at {bits} bits")
    p = gsim.generate("two-exponential-disks", 3000000, 33.0, 1.0, 10.0, seed=206)
    s = _sim(p, "barnes-hut")
    if bits != 10:
        s.set_tuning(gsim.TUNE_BH_LEAF_BITS, bits)
    s.step(1)
    t = s.tree()
    s.close()
    pp = p.padded()
    ot, nn, _, _, _ = O.bh_build(pp.pos, pp.mass, bits=bits)
    assert t["total"] == nn
    for k in ("sorted_src", "counts", "radii", "node_mass", "node_pos"):
        assert np.array_equal(t[k], ot[k]), k


@pytest.mark.parametrize("split", [1, 2])
def test_bh_heavy_list_walks_match_the_main_walk(split):
    """With more groups than the GPU holds at once (9,375 here), from the
    second step on the heaviest groups (at least `split` times the mean chunk
    count last step, at most 512) are walked by four waves each on a second
    stream beside the main launch: every step's result is bit-identical to one
    wave per group (heavy list off), and so are the counts."""
    _say(f"bh heavy list, split {split}")
    p = gsim.generate("two-exponential-disks", 600000, 33.0, 1.0, 10.0, seed=207)
    out = []
    for sp in (0, split):
        s = _sim(p, "barnes-hut")
        s.set_tuning(gsim.TUNE_BH_SPLIT, sp)
        s.set_tuning(gsim.TUNE_BH_ORDER, 0)  # the heavy list runs with index order only
        steps = []
        for _ in range(3):
            s.step(1)
            steps.append((s.dump(gsim.DUMP_ACCEL), s.dump(gsim.DUMP_COUNTERS)))
        g = s.read()
        s.close()
        out.append((steps, g))
    for (a0, c0), (a1, c1) in zip(out[0][0], out[1][0]):
        assert np.array_equal(a0, a1)
        assert np.array_equal(c0, c1)
    assert np.array_equal(out[0][1].pos, out[1][1].pos)
    assert np.array_equal(out[0][1].vel, out[1][1].vel)


def test_bh_leaf_sizes_at_chunk_boundaries():
    """Leaves of 24 (the largest summed by one lane), 25, 63, 64, 65, 127,
    128, 129, 200 and 1,000 particles: the wave-summed leaves mix full 64-node
    chunks and a ragged tail, and every sum must be the oracle's, bit for
    bit."""
    rng = np.random.default_rng(209)
    sizes = [24, 25, 63, 64, 65, 127, 128, 129, 200, 1000]
    pos, mass = [], []
    w = 1000.0 / 1024.0  # one depth-10 cell per leaf, well inside its edges
    for i, k in enumerate(sizes):
        corner = np.array([-500.0 + (100 + 80 * i) * w, -500.0 + 600 * w])
        pos.append((corner + rng.uniform(0.1, 0.85, size=(k, 2)) * w).astype(np.float32))
        mass.append(rng.uniform(1, 10, k).astype(np.float32))
    pos = np.vstack(pos).astype(np.float32)
    mass = np.concatenate(mass)
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = _sim(p, "barnes-hut")
    s.step(1)
    t = s.tree()
    s.close()
    pp = p.padded()
    ot, nn, *_ = O.bh_build(pp.pos, pp.mass)
    assert t["total"] == nn
    for k in ("sorted_src", "counts", "radii", "node_mass", "node_pos"):
        assert np.array_equal(t[k], ot[k]), k
    # every leaf is one bucket: its cell node sits right before its particles
    leaves = [i for i in range(nn) if t["radii"][i] == np.float32(500000.0 / 4.0**10)]
    assert sorted(int(t["counts"][i]) - 1 for i in leaves) == sorted(sizes + [len(pp.mass) - len(mass)])


def test_bh_degenerate_trees():
    cases = {
        "single": (np.array([[1.0, 2.0]], np.float32), np.array([3.0], np.float32)),
        "all-killed": (np.full((70, 2), 900.0, np.float32), np.ones(70, np.float32)),
        "one-leaf": (np.full((64, 2), 0.1, np.float32), np.ones(64, np.float32)),
        "edge": (np.array([[-500, -500], [499.99994, 499.99994], [499.99997, 0]],
                          np.float32), np.ones(3, np.float32)),
    }
    for name, (pos, mass) in cases.items():
        _say(f"bh degenerate {name}")
        p = gsim.Particles(pos, np.ones_like(pos), mass)
        s = _sim(p, "barnes-hut")
        s.step(1)
        t = s.tree()
        g = s.read()
        s.close()
        pp = p.padded()
        po, vo, mo, V, A = O.bh_step(pp.pos, pp.vel, pp.mass, 1.0, 1e-3, EPS2, 1.0, 64)
        ot, nn, *_ = O.bh_build(pp.pos, pp.mass)
        assert t["total"] == nn
        assert np.array_equal(t["counts"], ot["counts"])
        assert np.array_equal(t["sorted_src"], ot["sorted_src"])
        assert np.array_equal(g.mass, mo[: p.n])
        assert np.allclose(g.pos, po[: p.n], rtol=1e-6, atol=1e-6)
        assert np.allclose(g.vel, vo[: p.n], rtol=1e-5, atol=1e-5)


# ---- Barnes-Hut: traversal -----------------------------------------------------------------

@pytest.mark.parametrize("W", [64, 32])
@pytest.mark.parametrize("case", [1, 2, 3, 4])
def test_bh_traversal_counts_exact_forces_within_tol(W, case):
    """On the GPU's own (bit-exact) tree, the GPU walk visits and accepts the
    same nodes as the oracle's walk, so V/A match exactly; forces differ only
    by rsqrt rounding."""
    name, pos, mass = list(_bh_inputs())[case]
    _say(f"bh traverse W={W} {name}")
    vel = np.zeros_like(pos)
    p = gsim.Particles(pos, vel, mass)
    s = _sim(p, "barnes-hut", group_width=W)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)
    V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
    t = s.tree()
    g = s.read()
    s.close()
    pp = p.padded()
    _, _, _, pk, mk = O.bh_build(pp.pos, pp.mass)
    oa, oV, oA, amb, _, oa64 = O.bh_traverse(t, pk, mk, EPS2, 1.0, W=W, f64=True)
    _say(f"  V={V} A={A} oracle V={oV} A={oA}")
    assert (V, A) == (oV, oA)
    inside = t["sorted_src"][t["sorted_src"] < len(mk)]
    _bh_force_ok(a[inside], oa[inside], oa64[inside])
    # integrate of the in-domain particles
    kv = np.float32(1.0) * np.float32(1e-3)
    kp = np.float32(0.5) * np.float32(1e-3)
    real = inside[inside < p.n]
    nv = (pp.vel[real] + a[real] * kv).astype(np.float32)
    assert np.array_equal(g.vel[real], nv)
    assert np.array_equal(g.pos[real], (pk[real] + (pp.vel[real] + nv) * kp).astype(np.float32))


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 6])
@pytest.mark.parametrize("case", [2, 4])
def test_bh_traversal_variants_agree(mode, case):
    """Every traversal launch variant (static or persistent groups, scalar or
    LDS-staged leaf runs, cost-balanced segments) visits exactly the oracle's
    nodes."""
    name, pos, mass = list(_bh_inputs())[case]
    _say(f"bh traversal mode={mode} {name}")
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = _sim(p, "barnes-hut")
    s.set_tuning(gsim.TUNE_BH_TRAVERSAL, mode)
    s.set_tuning(gsim.TUNE_BH_WAVES_PER_SIMD, 2)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)
    V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
    t = s.tree()
    s.close()
    pp = p.padded()
    _, _, _, pk, mk = O.bh_build(pp.pos, pp.mass)
    oa, oV, oA, _, _, oa64 = O.bh_traverse(t, pk, mk, EPS2, 1.0, W=64, f64=True)
    assert (V, A) == (oV, oA)
    inside = t["sorted_src"][t["sorted_src"] < len(mk)]
    _bh_force_ok(a[inside], oa[inside], oa64[inside])


@pytest.mark.parametrize("segments,split", [(1, 3), (7, 3), (512, 0), (512, 1), (512, 3),
                                            (4096, 2)])
def test_bh_balanced_traversal_matches_static(segments, split):
    """The cost-balanced traversal (segments cut by the previous step's walk
    costs, dense-core groups split across four waves) gives bit for bit the
    static launch's results over several steps, for any number of segments,
    including more segments than groups, and any split threshold."""
    _say(f"bh balanced segments={segments} split={split}")
    p = gsim.generate("two-exponential-disks", 100000, 33.0, 1.0, 10.0, seed=208)
    out = []
    for mode in (2, 6):
        s = gsim.Simulation(p, "barnes-hut", accuracy_parameter=0.7)
        s.set_tuning(gsim.TUNE_BH_TRAVERSAL, mode)
        s.set_tuning(gsim.TUNE_BH_SEGMENTS, segments)
        s.set_tuning(gsim.TUNE_BH_SPLIT, split)
        s.step(4)
        out.append((s.read(), tuple(int(x) for x in s.dump(gsim.DUMP_COUNTERS))))
        s.close()
    (a, ca), (b, cb) = out
    assert ca == cb
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    assert np.array_equal(a.mass, b.mass)


def test_bh_theta_zero_equals_sorted_direct_sum():
    p = gsim.generate("random", 2000, 60.0, 1.0, 10.0, seed=206)
    s = _sim(p, "barnes-hut", accuracy_parameter=0.0)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)
    V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
    s.close()
    pp = p.padded()
    ref = O.ds_accel_f32(pp.pos, pp.mass, EPS2)
    ref64 = O.ds_accel_f64(pp.pos, pp.mass, EPS2)
    assert A == len(pp.mass) ** 2
    _force_vs_f64(a[: p.n], ref[: p.n], ref64[: p.n])


@pytest.mark.parametrize("W", [64, 32])
def test_bh_multistep_trajectory(W):
    """20 steps of a galaxy collision: the GPU run and the oracle run stay
    within FP32 rounding of each other (accelerations differ by rsqrt ulps,
    which can flip a vote near its threshold, so the bound is on the system
    norm of the trajectory)."""
    _say(f"bh multistep W={W}")
    p = gsim.generate("galaxy-collision", 6000, 150.0, 1.0, 10.0, seed=207)
    s = _sim(p, "barnes-hut", group_width=W, simulation_speed=2.0)
    s.step(20)
    g = s.read()
    s.close()
    pp = p.padded()
    po, vo, mo = pp.pos, pp.vel, pp.mass
    dt = float(np.float32(1e-3) * np.float32(2.0))
    for _ in range(20):
        po, vo, mo, _, _ = O.bh_step(po, vo, mo, 1.0, dt, EPS2, 1.0, W)
    disp = np.linalg.norm(po[: p.n] - p.pos)
    assert np.linalg.norm(g.pos - po[: p.n]) <= 1e-4 * disp
    assert H.system_rel_err(g.vel, vo[: p.n]) <= 1e-4
    assert np.array_equal(g.mass, mo[: p.n])


def test_timing_and_counters_api():
    p = gsim.generate("two-exponential-disks", 20000, 33.0, 1.0, 10.0, seed=208)
    s = gsim.Simulation(p, "barnes-hut")
    s.set_timing(True)
    s.step(3)
    t = s.timing()
    assert t["steps"] == 3 and t["force_ms"] > 0 and t["build_ms"] > 0
    assert abs(t["step_ms"] - t["force_ms"] - t["build_ms"]) < 1e-2 + 0.05 * t["step_ms"]
    s.close()


def test_w32_handle_refuses_w64_launch_knobs():
    """A group width 32 handle has one traversal launch shape: the W = 64
    launch knobs are refused instead of silently ignored."""
    p = gsim.generate("random", 2000, 60.0, 1.0, 10.0, seed=209)
    knobs = (gsim.TUNE_BH_WAVES_PER_SIMD, gsim.TUNE_BH_TRAVERSAL, gsim.TUNE_BH_SPLIT,
             gsim.TUNE_BH_SEGMENTS)
    with gsim.Simulation(p, "barnes-hut", group_width=32) as s:
        for knob in knobs:
            with pytest.raises(gsim.GsimError) as e:
                s.set_tuning(knob, 2)
            assert e.value.status == 1
        s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 0)  # not a launch knob
    with gsim.Simulation(p, "barnes-hut", group_width=64) as s:
        for knob in knobs:
            s.set_tuning(knob, 2)


@pytest.mark.parametrize("W", [32, 64])
def test_bh_xcd_group_order_is_bit_identical(W):
    """Handing each XCD runs of consecutive workgroups changes only where a
    group runs: states and visit/accept counts are bit-identical over steps,
    for run lengths that do and do not divide the launch."""
    p = gsim.generate("galaxy-collision", 40000, 120.0, 1.0, 10.0, seed=210)
    out = []
    for order in (1, 3, 16, 4096):
        with gsim.Simulation(p, "barnes-hut", group_width=W, accuracy_parameter=0.8) as s:
            s.set_tuning(gsim.TUNE_BH_XCD_ORDER, order)
            s.step(3)
            out.append((s.read(), tuple(int(x) for x in s.dump(gsim.DUMP_COUNTERS))))
    a, ca = out[0]
    for b, cb in out[1:]:
        assert ca == cb
        assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)


@pytest.mark.parametrize("W", [32, 64])
def test_bh_heaviest_first_order_is_bit_identical(W):
    """The static walk's workgroup order -- index, heaviest first by last
    step's visits (re-planned every 8 steps), and auto (a timed trial of both
    over steps 1-4), combined with XCD runs -- changes only which wave walks
    a group: states and visit/accept counts are bit-identical over steps
    (the first step has no costs yet and runs in index order)."""
    p = gsim.generate("galaxy-collision", 40000, 120.0, 1.0, 10.0, seed=211)
    out = []
    for order, xcd in ((0, 1), (1, 1), (2, 1), (1, 16)):
        with gsim.Simulation(p, "barnes-hut", group_width=W, accuracy_parameter=0.8) as s:
            s.set_tuning(gsim.TUNE_BH_ORDER, order)
            s.set_tuning(gsim.TUNE_BH_XCD_ORDER, xcd)
            s.step(5)
            s.synchronize()
            s.step(5)  # auto: the trial has completed; its pick walks these
            out.append((s.read(), tuple(int(x) for x in s.dump(gsim.DUMP_COUNTERS))))
    a, ca = out[0]
    for b, cb in out[1:]:
        assert ca == cb
        assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    with gsim.Simulation(p, "barnes-hut", group_width=32) as s:
        with pytest.raises(gsim.GsimError):
            s.set_tuning(gsim.TUNE_BH_ORDER, 3)


@pytest.mark.parametrize("W,bits", [(32, 10), (64, 10), (32, 15)])
def test_bh_walk_written_cells_match_positions(W, bits):
    """From the second step on the walk writes the next build's sort input
    (each particle's cell at its caller index) from the positions it
    integrates, and the build only counts digits.  The cells that build
    sorted equal the cells of the positions read back before it, the 30-bit
    codes of those positions extend them, and the sorted order maps back to
    the caller's indices as the oracle's build of the same positions does."""
    p = gsim.generate("two-exponential-disks", 20000, 33.0, 1.0, 10.0, 1.0, seed=212)
    with gsim.Simulation(p, "barnes-hut", group_width=W) as s:
        if bits != 10:
            s.set_tuning(gsim.TUNE_BH_LEAF_BITS, bits)
        s.step(3)
        st = s.read()
        s.step(1)
        cells = s.dump(gsim.DUMP_CELLS)[: p.n]
        if bits == 10:
            m30 = s.dump(gsim.DUMP_MORTON30)[: p.n]
            t = s.tree()
    if bits == 10:
        assert np.array_equal(cells, H.cells_numpy(st.pos))
        ok = cells != 0xFFFFFFFF
        assert np.array_equal(m30[ok] >> np.uint32(10), cells[ok])
        src = t["sorted_src"]
        real = src[src < p.n]
        # the sorted order by cell, ties by caller index, over the real particles
        order = np.lexsort((np.arange(p.n), cells))
        order = order[cells[order] != 0xFFFFFFFF]
        assert np.array_equal(real, order)
    else:
        want = np.array([O.cell_bits(float(x), float(y), None, 15) for x, y in st.pos], np.uint32)
        assert np.array_equal(cells, want)
