"""15-bit leaves (GSIM_TUNE_BH_LEAF_BITS = 15; beyond the reference, opt-in):
leaves on a 32768^2 grid at depth 15, alone and with the adaptive box.  The
30-bit keys, the sorted sources and every node array are compared bit for bit
with the oracle's sparse build at 15 bits; visits/acceptances exactly; forces
against the FP64 sum of the same accepted terms.  Switching back to 10 bits
gives the reference's tree again."""
import sys

import numpy as np
import pytest

import gsim
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

EPS2 = H.sq32(0.2)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inputs():
    rng = np.random.default_rng(501)
    pos = rng.normal(0, 200, size=(6000, 2)).astype(np.float32)
    pos[:2500] = rng.normal(30, 0.1, size=(2500, 2))  # a knot: one depth-10 bucket
    pos[2500:2506] = [[700, 0], [0, -900], [np.nan, 1], [3, 3], [3, 3], [-0.0, 0.0]]
    yield "knot-6000", pos, rng.uniform(1, 10, 6000).astype(np.float32)
    p = gsim.generate("galaxy-collision", 20000, 150.0, 1.0, 10.0, seed=502)
    yield "collision-150-20000", p.pos, p.mass
    yield "one-point-700", np.full((700, 2), 12.5, np.float32), np.ones(700, np.float32)
    p = gsim.generate("random", 3000, 0.05, 1.0, 10.0, seed=503)  # below one depth-10 cell
    yield "tiny-3000", p.pos, p.mass


def _sim(p, boxed, theta=1.0, bits=15, W=64):
    s = gsim.Simulation(p, "barnes-hut", accuracy_parameter=theta, group_width=W)
    s.set_tuning(gsim.TUNE_BH_LEAF_BITS, bits)
    if boxed:
        s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
    return s


@pytest.mark.parametrize("boxed", [False, True])
@pytest.mark.parametrize("case", range(4))
def test_leaf15_tree_bit_exact(case, boxed):
    name, pos, mass = list(_inputs())[case]
    _say(f"15-bit tree {name} box={boxed}")
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = _sim(p, boxed)
    s.step(1)
    t = s.tree()
    cells = s.dump(gsim.DUMP_CELLS)
    box = s.dump(gsim.DUMP_BOX) if boxed else None
    s.close()
    pp = p.padded()
    if boxed:
        assert np.array_equal(box, O.bh_box(pp.pos))
    want = np.array([O.cell_bits(float(x), float(y), box, 15) for x, y in pp.pos], np.uint32)
    assert np.array_equal(cells, want)
    ot, nn, nin, _, _ = O.bh_build(pp.pos, pp.mass, box=box, bits=15)
    assert t["total"] == nn
    for k in ("sorted_src", "counts", "radii", "node_mass", "node_pos"):
        assert np.array_equal(t[k], ot[k]), k


@pytest.mark.parametrize("W", [64, 32])
@pytest.mark.parametrize("boxed", [False, True])
@pytest.mark.parametrize("case", [0, 1])
def test_leaf15_traversal_counts_and_forces(case, boxed, W):
    name, pos, mass = list(_inputs())[case]
    _say(f"15-bit traversal {name} box={boxed} W={W}")
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = _sim(p, boxed, theta=0.7, W=W)
    s.set_debug(gsim.DEBUG_ACCEL)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)
    V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
    t = s.tree()
    box = s.dump(gsim.DUMP_BOX) if boxed else None
    s.close()
    pp = p.padded()
    _, _, _, pk, mk = O.bh_build(pp.pos, pp.mass, box=box, bits=15)
    oa, oV, oA, _, _, oa64 = O.bh_traverse(t, pk, mk, EPS2, H.sq32(0.7), W=W, f64=True)
    assert (V, A) == (oV, oA)
    inside = t["sorted_src"][t["sorted_src"] < len(mk)]
    assert H.system_rel_err(a[inside], oa[inside]) <= 1e-5
    eg = H.per_particle_rel(a[inside], oa64[inside])
    e32 = H.per_particle_rel(oa[inside], oa64[inside])
    assert eg.max() <= max(1.5 * e32.max(), 1e-6)


def test_leaf15_splits_the_knot_and_switches_back():
    """The knot is one depth-10 bucket that every group near it streams whole
    (exact, and quadratic in the knot's size); at 15 bits groups accept its
    sub-cells, so far fewer terms are summed and the error is the Barnes-Hut
    error set by theta: it falls with theta.  Setting 10 bits again restores
    the reference's tree bit for bit."""
    name, pos, mass = next(_inputs())
    n = len(mass) // 64 * 64  # no padding: read() then returns the whole state
    p = gsim.Particles(pos[:n], np.zeros_like(pos[:n]), mass[:n])
    pp = p.padded()
    _, _, _, pk, mk = O.bh_build(pp.pos, pp.mass)  # the kills of the fixed domain
    ref = O.ds_accel_f64(pk, mk, EPS2)
    live = mk > 0
    out = {}
    for bits, theta in ((10, 1.0), (15, 1.0), (15, 0.5), (15, 0.25)):
        s = _sim(p, False, theta=theta, bits=bits)
        s.set_debug(gsim.DEBUG_ACCEL)
        s.step(1)
        a = s.dump(gsim.DUMP_ACCEL)
        _, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
        s.close()
        e = H.per_particle_rel(a[live], ref[live])
        out[bits, theta] = (A, float(np.quantile(e, 0.99)), float(e.max()))
        _say(f"  {bits}-bit leaves, theta {theta}: {A} accepted terms, error p99 "
             f"{out[bits, theta][1]:.2e} max {out[bits, theta][2]:.2e}")
    assert out[10, 1.0][2] < 1e-4                       # the bucket is summed exactly
    assert out[15, 1.0][0] * 4 < out[10, 1.0][0]        # far fewer terms
    assert out[15, 0.25][0] < out[10, 1.0][0]
    assert out[15, 1.0][1] > out[15, 0.5][1] > out[15, 0.25][1]
    assert out[15, 0.25][2] < 2e-2
    s = _sim(p, False, bits=15)
    s.step(1)
    g = s.read()  # the second step's input
    s.set_tuning(gsim.TUNE_BH_LEAF_BITS, 10)
    s.step(1)
    t = s.tree()
    s.close()
    ot, nn, _, _, _ = O.bh_build(g.pos, g.mass)
    assert t["total"] == nn
    for k in ("sorted_src", "counts", "radii", "node_mass", "node_pos"):
        assert np.array_equal(t[k], ot[k]), k


def test_leaf15_knob_errors():
    p = gsim.generate("random", 1000, 50.0, 1.0, 10.0, seed=504)
    s = gsim.Simulation(p, "barnes-hut")
    with pytest.raises(gsim.GsimError):
        s.set_tuning(gsim.TUNE_BH_LEAF_BITS, 12)
    s.close()
    d = gsim.Simulation(p, "direct-sum")
    with pytest.raises(gsim.GsimError):
        d.set_tuning(gsim.TUNE_BH_LEAF_BITS, 15)
    d.close()


def test_leaf15_steps_conserve_mass_and_momentum():
    """Twenty 15-bit steps of a collision inside the box: no particle is lost
    and the total momentum drifts no more than with 10-bit leaves."""
    p = gsim.generate("galaxy-collision", 20000, 150.0, 1.0, 10.0, seed=505)
    drift = {}
    for bits in (10, 15):
        s = _sim(p, True, bits=bits)
        s.step(20)
        g = s.read()
        s.close()
        assert np.array_equal(g.mass, p.mass) and np.isfinite(g.pos).all()
        m = p.mass.astype(np.float64)[:, None]
        mom0 = (m * p.vel).sum(0)
        mom = (m * g.vel.astype(np.float64)).sum(0)
        drift[bits] = float(np.linalg.norm(mom - mom0) / (np.abs(m * p.vel).sum()))
        _say(f"  {bits}-bit leaves: momentum drift {drift[bits]:.2e}")
    assert drift[15] <= max(4 * drift[10], 1e-6)
