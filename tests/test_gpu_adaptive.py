"""Adaptive bounding box (GSIM_TUNE_BH_ADAPTIVE_BOX; beyond the reference,
opt-in): the tree spans the particles' bounding square instead of the fixed
[-500, 500) domain, so no finite particle is killed.  The box, the cells and
every node array are compared bit for bit with the oracle's build over the
same box; visits/acceptances exactly; forces against the FP64 sum of the same
accepted terms; and the force error against the FP64 direct sum (the
theta-error) is reported and must fall with theta."""
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
    rng = np.random.default_rng(401)
    pos = rng.normal(0, 400, size=(5000, 2)).astype(np.float32)
    pos[:4] = [[2600, 30], [-1700, -2100], [np.nan, 3], [900, 0]]
    yield "wide-5000", pos, rng.uniform(1, 10, 5000).astype(np.float32)
    p = gsim.generate("galaxy-collision", 20000, 600.0, 1.0, 10.0, seed=402)
    yield "collision-600-20000", p.pos, p.mass
    p = gsim.generate("random", 3000, 0.05, 1.0, 10.0, seed=403)  # a tiny box
    yield "tiny-3000", p.pos, p.mass
    yield "one-point-700", np.full((700, 2), 12.5, np.float32), np.ones(700, np.float32)


@pytest.mark.parametrize("case", range(4))
def test_adaptive_tree_bit_exact(case):
    name, pos, mass = list(_inputs())[case]
    _say(f"adaptive tree {name}")
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = gsim.Simulation(p, "barnes-hut")
    s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
    s.step(1)
    box = s.dump(gsim.DUMP_BOX)
    t = s.tree()
    cells = s.dump(gsim.DUMP_CELLS)
    with pytest.raises(gsim.GsimError):
        s.dump(gsim.DUMP_MORTON30)  # codes of the fixed domain only
    s.close()
    pp = p.padded()
    obox = O.bh_box(pp.pos)
    assert np.array_equal(box, obox) and np.array_equal(obox, H.box_numpy(pp.pos))
    ot, nn, nin, pk, mk = O.bh_build(pp.pos, pp.mass, box=obox)
    ocells = np.array([O.cell20_box(float(x), float(y), obox) for x, y in pp.pos], np.uint32)
    assert np.array_equal(cells, ocells)
    assert nin == int(np.isfinite(pp.pos).all(1).sum())  # only non-finite positions leave
    assert t["total"] == nn
    for k in ("sorted_src", "counts", "radii", "node_mass", "node_pos"):
        assert np.array_equal(t[k], ot[k]), k


@pytest.mark.parametrize("W", [64, 32])
@pytest.mark.parametrize("case", [0, 1])
def test_adaptive_traversal_counts_and_forces(case, W):
    name, pos, mass = list(_inputs())[case]
    _say(f"adaptive traversal {name} W={W}")
    p = gsim.Particles(pos, np.zeros_like(pos), mass)
    s = gsim.Simulation(p, "barnes-hut", accuracy_parameter=0.7, group_width=W)
    s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
    s.set_debug(gsim.DEBUG_ACCEL)
    s.step(1)
    a = s.dump(gsim.DUMP_ACCEL)
    V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
    t = s.tree()
    s.close()
    pp = p.padded()
    box = O.bh_box(pp.pos)
    _, _, _, pk, mk = O.bh_build(pp.pos, pp.mass, box=box)
    oa, oV, oA, _, _, oa64 = O.bh_traverse(t, pk, mk, EPS2, H.sq32(0.7), W=W, f64=True)
    assert (V, A) == (oV, oA)
    inside = t["sorted_src"][t["sorted_src"] < len(mk)]
    assert H.system_rel_err(a[inside], oa[inside]) <= 1e-5
    eg = H.per_particle_rel(a[inside], oa64[inside])
    e32 = H.per_particle_rel(oa[inside], oa64[inside])
    assert eg.max() <= max(1.5 * e32.max(), 1e-6)


def test_adaptive_box_keeps_particles_and_theta_error_falls():
    """A galaxy collision of size 600: the fixed domain kills the particles
    beyond +-500 in the first build; the adaptive box keeps every one over 20
    steps.  Force error against the FP64 direct sum (median / p99 / max per
    particle, relative to max(|a|, rms)) falls with theta."""
    p = gsim.generate("galaxy-collision", 20000, 600.0, 1.0, 10.0, seed=404)
    out = p.pos.copy()
    far = (np.abs(out) >= 500).any(1).sum()
    assert far > 100
    fixed = gsim.Simulation(p, "barnes-hut")
    fixed.step(1)
    assert (fixed.read().mass == 0).sum() >= far
    fixed.close()
    s = gsim.Simulation(p, "barnes-hut")
    s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
    s.step(20)
    g = s.read()
    s.close()
    assert np.array_equal(g.mass, p.mass) and np.isfinite(g.pos).all()
    pp = p.padded()
    ref = O.ds_accel_f64(pp.pos, pp.mass, EPS2)[: p.n]
    errs = []
    for theta in (1.0, 0.5, 0.25):
        s = gsim.Simulation(p, "barnes-hut", accuracy_parameter=theta)
        s.set_tuning(gsim.TUNE_BH_ADAPTIVE_BOX, 1)
        s.set_debug(gsim.DEBUG_ACCEL)
        s.step(1)
        a = s.dump(gsim.DUMP_ACCEL)[: p.n]
        s.close()
        e = H.per_particle_rel(a, ref)
        errs.append((theta, float(np.median(e)), float(np.quantile(e, 0.99)), float(e.max())))
        _say(f"  theta {theta}: median {errs[-1][1]:.2e} p99 {errs[-1][2]:.2e} max {errs[-1][3]:.2e}")
    assert errs[0][1] > errs[1][1] > errs[2][1]
    assert errs[2][2] < 1e-2
