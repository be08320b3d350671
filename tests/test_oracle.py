"""Pins the CPU oracle (the checker) against closed forms, brute force and
invariants — none of these retype the oracle's own formula.  CPU only."""
import numpy as np
import pytest

import helpers as H
import oracle as O


# ---- Morton cells (InitShader.comp:80-98) --------------------------------------

def test_morton_exhaustive_against_bit_loop():
    """All 2^20 (ix, iy): the oracle's cell of the centre of leaf (ix, iy)
    equals the bit-loop interleave; 2^20 C calls are too slow via ctypes, so the
    mask form is checked in numpy against the bit loop exhaustively and the
    oracle on a dense sample of the same centres."""
    ix, iy = np.meshgrid(np.arange(1024, dtype=np.uint32), np.arange(1024, dtype=np.uint32))
    ix = ix.ravel(); iy = iy.ravel()
    ref = H.morton_bitloop(ix, iy)

    def part1by1(v):
        v = v.astype(np.uint32)
        v = (v | (v << 8)) & np.uint32(0x00FF00FF)
        v = (v | (v << 4)) & np.uint32(0x0F0F0F0F)
        v = (v | (v << 2)) & np.uint32(0x33333333)
        v = (v | (v << 1)) & np.uint32(0x55555555)
        return v

    assert np.array_equal(part1by1(ix) | (part1by1(iy) << np.uint32(1)), ref)
    # positions at leaf centres: x = (ix + 0.5) * (1000/1024) - 500
    rng = np.random.default_rng(0)
    sel = rng.choice(ix.size, 4000, replace=False)
    xs = ((ix[sel] + 0.5) * (1000.0 / 1024.0) - 500.0).astype(np.float32)
    ys = ((iy[sel] + 0.5) * (1000.0 / 1024.0) - 500.0).astype(np.float32)
    got = np.array([O.cell20(float(x), float(y)) for x, y in zip(xs, ys)], np.uint32)
    assert np.array_equal(got, ref[sel])


@pytest.mark.parametrize("x,y,expect_inside", [
    (-500.0, -500.0, True),       # lower edge is inside
    (500.0, 0.0, False),          # q = 1024 is outside
    (0.0, 500.0, False),
    # one ulp below 500: x/500 + 1 ties to 2.0, so q = 1024 and the particle
    # is killed by the reference formula; two ulps below stays inside
    (np.nextafter(np.float32(500), np.float32(0)), 0.0, False),
    (np.float32(500) - np.float32(2.0 ** -14), 0.0, True),
    (-500.0001, 0.0, False),
    (float("nan"), 0.0, False),   # NaN fails every comparison -> killed
    (0.0, float("inf"), False),
    (0.0, 0.0, True),
])
def test_cell_boundaries(x, y, expect_inside):
    c = O.cell20(float(np.float32(x)), float(np.float32(y)))
    assert (c != 0xFFFFFFFF) == expect_inside


def test_cell_known_answers():
    # origin: q = 512 on both axes -> part1by1(512) | part1by1(512) << 1
    assert O.cell20(0.0, 0.0) == (1 << 18) | (1 << 19)
    assert O.cell20(-500.0, -500.0) == 0
    last = np.float32(500) - np.float32(2.0 ** -14)
    assert O.cell20(float(last), float(last)) == (1 << 20) - 1
    # the numpy FP32 restatement agrees on random positions, inside and out
    rng = np.random.default_rng(1)
    pos = rng.uniform(-520, 520, size=(5000, 2)).astype(np.float32)
    got = np.array([O.cell20(float(a), float(b)) for a, b in pos], np.uint32)
    assert np.array_equal(got, H.cells_numpy(pos))


def test_morton30_prefix_is_cell20():
    """30-bit code of the same normalised coordinate, shifted right by 10, is
    the 20-bit cell (scaling by 32 is exact)."""
    rng = np.random.default_rng(2)
    pos = rng.uniform(-499.9, 499.9, size=(20000, 2)).astype(np.float32)
    q = ((pos / np.float32(500)) + np.float32(1)) * np.float32(0.5)
    q30 = (q * np.float32(32768)).astype(np.uint32)
    m30 = H.morton_bitloop(q30[:, 0], q30[:, 1], bits=15)
    assert np.array_equal(m30 >> np.uint32(10), H.cells_numpy(pos))


# ---- direct sum (SimShader.comp) -----------------------------------------------

def test_ds_two_body_closed_form():
    pos = np.array([[0.0, 0.0], [3.0, 4.0]], np.float32)
    mass = np.array([2.0, 5.0], np.float32)
    eps2 = H.sq32(0.2)
    a = O.ds_accel_f64(pos, mass, eps2)
    r2 = 25.0 + eps2
    k = 1.0 / (r2 * np.sqrt(r2))
    assert np.allclose(a[0], np.array([3.0, 4.0]) * 5.0 * k, rtol=1e-12)
    assert np.allclose(a[1], -np.array([3.0, 4.0]) * 2.0 * k, rtol=1e-12)
    a32 = O.ds_accel_f32(pos, mass, eps2)
    assert np.allclose(a32, a, rtol=1e-6)


def test_ds_ring_centre_has_zero_force():
    n = 64
    ang = 2 * np.pi * np.arange(n) / n
    pos = np.zeros((n + 1, 2), np.float32)
    pos[1:, 0] = 10 * np.cos(ang)
    pos[1:, 1] = 10 * np.sin(ang)
    mass = np.ones(n + 1, np.float32)
    a64 = O.ds_accel_f64(pos, mass, H.sq32(0.2))
    ring = np.linalg.norm(a64[1:], axis=1).mean()
    assert np.linalg.norm(a64[0]) < 1e-12 * ring * n


def test_ds_newton_third_law_and_f64_agreement():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-50, 50, size=(512, 2)).astype(np.float32)
    mass = rng.uniform(1, 10, size=512).astype(np.float32)
    eps2 = H.sq32(0.2)
    a64 = O.ds_accel_f64(pos, mass, eps2)
    mom = (mass[:, None].astype(np.float64) * a64).sum(axis=0)
    scale = np.abs(mass[:, None] * a64).sum()
    assert np.linalg.norm(mom) < 1e-12 * scale
    a32 = O.ds_accel_f32(pos, mass, eps2)
    assert H.system_rel_err(a32, a64) < 2e-6


def test_ds_accel_f64_at_equals_full_rows():
    rng = np.random.default_rng(15)
    pos = rng.uniform(-30, 30, size=(700, 2)).astype(np.float32)
    mass = rng.uniform(1, 10, size=700).astype(np.float32)
    full = O.ds_accel_f64(pos, mass, 0.04)
    targets = np.array([699, 0, 5, 5, 321, 64], np.uint32)
    at = O.ds_accel_f64_at(pos, mass, 0.04, targets, nthreads=3)
    assert np.array_equal(at, full[targets])


def test_ds_step_free_particle_and_thread_invariance():
    pos = np.array([[1.0, 2.0]], np.float32)
    vel = np.array([[3.0, -4.0]], np.float32)
    mass = np.array([7.0], np.float32)
    po, vo = O.ds_step(pos, vel, mass, 1.0, 1e-3, H.sq32(0.2))
    # self term: d = 0 -> no force; p' = p + ((v + v)*0.5)*dt
    assert np.array_equal(vo, vel)
    dt = np.float32(1e-3)
    assert np.array_equal(po, pos + ((vel + vel) * np.float32(0.5)) * dt)
    rng = np.random.default_rng(4)
    p = rng.uniform(-20, 20, size=(300, 2)).astype(np.float32)
    m = rng.uniform(1, 3, size=300).astype(np.float32)
    a1 = O.ds_accel_f32(p, m, 0.04, nthreads=1)
    a7 = O.ds_accel_f32(p, m, 0.04, nthreads=7)
    assert np.array_equal(a1, a7)


def test_ds_zero_softening_gives_nan_like_reference():
    pos = np.array([[0.0, 0.0], [1.0, 0.0]], np.float32)
    a = O.ds_accel_f32(pos, np.ones(2, np.float32), 0.0)
    assert np.isnan(a).all()  # self term 0 * inf


# ---- direct-sum steps: the FP32 step and the FP64 trajectory yardstick -----------
# SimShader.comp:53-69: a = sum_j m_j d (d.d + eps^2)^(-3/2); v' = v + a (dt G);
# p' = p + ((v + v') / 2) dt.  The expected values below come from Newton's law
# written out in numpy (float64), from conservation laws and from the orbit's
# geometry, never from the oracle's code.

def _two_body_expect(pos, mass, vel, G, dt, eps2):
    """One step of two bodies in float64, written from the physics: each body
    is pulled toward the other by G m_other d / (|d|^2 + eps^2)^(3/2)."""
    d = pos[1].astype(np.float64) - pos[0].astype(np.float64)
    k = 1.0 / (d @ d + eps2) ** 1.5
    acc = np.array([mass[1] * d * k, -mass[0] * d * k]) * G
    v1 = vel.astype(np.float64) + acc * dt
    p1 = pos.astype(np.float64) + (vel.astype(np.float64) + v1) * 0.5 * dt
    return p1, v1


def test_ds_steps_f64_free_particle_is_exact():
    pos = np.array([[1.25, -2.5]], np.float64)
    vel = np.array([[3.0, -4.0]], np.float64)
    for steps in (1, 7):
        p, v = O.ds_steps_f64(pos, vel, np.array([7.0]), 0.9, 1e-3, 0.04, steps)
        assert np.array_equal(v, vel)
        expect = pos.copy()
        for _ in range(steps):
            expect = expect + vel * 1e-3  # (v + v) / 2 == v exactly
        assert np.array_equal(p, expect)


@pytest.mark.parametrize("G,eps", [(1.0, 0.2), (0.7, 1.5), (2.5, 0.01)])
def test_ds_steps_f64_two_bodies_from_rest(G, eps):
    """Unequal masses (a mass-index swap fails), G != 1 (a dropped or misplaced
    G fails), softening comparable to the separation (a missing or unsquared
    eps fails), and the half-step position update (p' = p + v' dt / 2 from
    rest: a full-step or missing update fails)."""
    pos = np.array([[0.5, -1.0], [2.0, 1.0]])
    mass = np.array([2.0, 5.0])
    vel = np.zeros((2, 2))
    dt, eps2 = 1e-2, eps * eps
    p, v = O.ds_steps_f64(pos, vel, mass, G, dt, eps2, 1)
    pe, ve = _two_body_expect(pos, mass, vel, G, dt, eps2)
    assert np.allclose(v, ve, rtol=1e-14, atol=0)
    assert np.allclose(p, pe, rtol=1e-14, atol=0)
    # the pull points from each body toward the other
    assert np.dot(v[0], pos[1] - pos[0]) > 0 and np.dot(v[1], pos[0] - pos[1]) > 0


def test_ds_steps_f64_two_moving_bodies_two_steps():
    """Two steps with nonzero velocities: the second step must use the first
    step's positions and velocities."""
    pos = np.array([[-1.0, 0.25], [1.5, -0.5]])
    vel = np.array([[0.3, -0.2], [-0.1, 0.4]])
    mass = np.array([3.0, 1.5])
    G, dt, eps2 = 1.3, 5e-3, 0.09
    p, v = O.ds_steps_f64(pos, vel, mass, G, dt, eps2, 2)
    pe, ve = _two_body_expect(pos, mass, vel, G, dt, eps2)
    pe, ve = _two_body_expect(pe, mass, ve, G, dt, eps2)
    assert np.allclose(v, ve, rtol=1e-13, atol=0)
    assert np.allclose(p, pe, rtol=1e-13, atol=0)


def test_ds_steps_f64_conserves_momentum_and_com_motion():
    """64 random bodies, 1,000 steps: forces are pairwise equal and opposite,
    so sum m v is constant and the centre of mass moves uniformly
    (sum m p' = sum m p + dt (P + P') / 2 = sum m p + dt P) to rounding.  A
    wrong mass index or an unsymmetric softening breaks both."""
    rng = np.random.default_rng(11)
    n = 64
    pos = rng.uniform(-10, 10, size=(n, 2))
    vel = rng.normal(0, 1, size=(n, 2))
    mass = rng.uniform(1, 10, size=n)
    G, dt, eps2, steps = 1.0, 1e-3, 0.04, 1000
    p, v = O.ds_steps_f64(pos, vel, mass, G, dt, eps2, steps)
    P0 = (mass[:, None] * vel).sum(0)
    P1 = (mass[:, None] * v).sum(0)
    scale = (mass[:, None] * np.abs(v)).sum()
    assert np.abs(P1 - P0).max() < 1e-13 * scale * 10
    com0 = (mass[:, None] * pos).sum(0) / mass.sum()
    com1 = (mass[:, None] * p).sum(0) / mass.sum()
    assert np.allclose(com1, com0 + P0 / mass.sum() * dt * steps, rtol=0, atol=1e-10)
    # and the bodies did interact: velocities changed by far more than rounding
    assert np.abs(v - vel).max() > 1e-2


def test_ds_steps_f64_circular_orbit():
    """Two equal masses on the softened circular orbit: separation D, speed
    u = sqrt(G m D^2 / (2 (D^2 + eps^2)^(3/2))) each, period T = pi D / u.
    After one period the separation is still D and each body is back at its
    start, to the integrator's error, which falls linearly with dt (its
    velocity update is first order)."""
    G, m, D, eps2 = 1.0, 4.0, 2.0, 0.04
    u = np.sqrt(G * m * D * D / (2.0 * (D * D + eps2) ** 1.5))
    T = np.pi * D / u
    drift, errs = [], []
    for nsteps in (2000, 4000):
        dt = T / nsteps
        pos = np.array([[-D / 2, 0.0], [D / 2, 0.0]])
        vel = np.array([[0.0, -u], [0.0, u]])
        p, v = O.ds_steps_f64(pos, vel, np.array([m, m]), G, dt, eps2, nsteps)
        drift.append(abs(np.linalg.norm(p[1] - p[0]) - D))
        errs.append(np.abs(p - pos).max())
    assert drift[0] < 0.025 * D and errs[0] < 0.05 * D
    # first-order convergence: halving dt halves both errors
    assert 1.8 < drift[0] / drift[1] < 2.2
    assert 1.8 < errs[0] / errs[1] < 2.2


def test_ds_step_f32_two_bodies_within_rounding():
    """The FP32 step against the float64 physics: within a few FP32 ulp of the
    velocity change and of the position update."""
    pos = np.array([[0.5, -1.0], [2.0, 1.0]], np.float32)
    vel = np.array([[0.25, 0.5], [-0.75, 0.125]], np.float32)
    mass = np.array([2.0, 5.0], np.float32)
    G, dt, eps2 = 0.7, 1e-2, H.sq32(0.2)
    po, vo = O.ds_step(pos, vel, mass, G, dt, eps2)
    # the oracle rounds dt and G to FP32 as the shader's constants are
    pe, ve = _two_body_expect(pos, mass, vel, float(np.float32(G)), float(np.float32(dt)), eps2)
    dv = ve - vel
    assert np.allclose(vo - vel, dv, rtol=1e-5, atol=0)
    assert np.allclose(po, pe, rtol=0, atol=4 * np.finfo(np.float32).eps * np.abs(pe).max())


def test_ds_step_f32_tracks_f64_trajectory():
    """The FP32 step and the FP64 yardstick agree to FP32 rounding over a few
    steps of a 128-body cluster, and the FP64 run is not the FP32 one."""
    rng = np.random.default_rng(12)
    n = 128
    pos = rng.uniform(-20, 20, size=(n, 2)).astype(np.float32)
    vel = rng.normal(0, 0.5, size=(n, 2)).astype(np.float32)
    mass = rng.uniform(1, 10, size=n).astype(np.float32)
    G, dt, eps2 = 1.0, 1e-3, H.sq32(0.2)
    p32, v32 = pos, vel
    for _ in range(5):
        p32, v32 = O.ds_step(p32, v32, mass, G, dt, eps2)
    p64, v64 = O.ds_steps_f64(pos, vel, mass, G, dt, eps2, 5)
    disp = np.abs(p64 - pos).max()
    assert np.abs(p32 - p64).max() < 1e-4 * disp + 8 * np.finfo(np.float32).eps * 20
    assert np.abs(v32 - v64).max() < 1e-5 * np.abs(v64 - vel).max() + 1e-6
    assert np.abs(v64 - vel).max() > 1e-3


# ---- Barnes-Hut build ------------------------------------------------------------

def _compare_tree(t, b):
    assert t["total"] == b["total"]
    assert np.array_equal(t["counts"], b["counts"])
    assert np.array_equal(t["radii"], b["radii"])
    assert np.array_equal(t["node_pos"], b["node_pos"])
    assert np.array_equal(t["node_mass"], b["node_mass"])
    assert np.array_equal(t["sorted_src"], b["sorted_src"])


@pytest.mark.parametrize("seed,n,spread", [(0, 37, 400.0), (1, 64, 5.0), (2, 150, 1.0),
                                           (3, 200, 0.01)])
def test_bh_build_matches_recursive_quadtree(seed, n, spread):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-spread, spread, size=(n, 2)).astype(np.float32)
    mass = rng.uniform(1, 10, size=n).astype(np.float32)
    pad = (-n) % 64
    pos = np.vstack([pos, np.zeros((pad, 2), np.float32)])
    mass = np.concatenate([mass, np.zeros(pad, np.float32)])
    tree, nn, nin, _, _ = O.bh_build(pos, mass)
    _compare_tree(tree, H.brute_tree(pos, mass))


def test_bh_build_counts_invariants():
    rng = np.random.default_rng(5)
    pos = np.vstack([rng.normal(0, 3, size=(900, 2)), rng.uniform(-480, 480, size=(124, 2))])
    pos = pos.astype(np.float32)
    mass = rng.uniform(1, 10, size=1024).astype(np.float32)
    tree, nn, nin, _, _ = O.bh_build(pos, mass)
    cells = H.cells_numpy(pos)
    inside = cells[cells != 0xFFFFFFFF]
    # counts[0] = N_in + #cells with n >= 2 over all depths
    ncell = 0
    for d in range(11):
        _, cnt = np.unique(inside >> np.uint32(2 * (10 - d)), return_counts=True)
        ncell += int((cnt >= 2).sum())
    assert nin == inside.size
    assert tree["counts"][0] == nin + ncell == nn
    # every subtree ends inside its parent's subtree (pre-order consistency)
    c = tree["counts"]
    for u in range(nn):
        end = u + c[u]
        assert end <= nn
        v = u + 1
        while v < end:
            v += c[v]
        assert v == end
    # particle nodes are exactly the radius-0 nodes, one per inside particle
    assert int((tree["radii"] == 0).sum()) == nin
    assert np.all(tree["counts"][tree["radii"] == 0] == 1)


def test_bh_build_dyadic_com_exact_and_kills():
    pos, vel, mass = H.dyadic_particles(700, seed=6)
    pos[:5] = [[600, 0], [0, -700], [float("nan"), 1], [500, 500], [-500, -500]]
    pos = np.vstack([pos, np.zeros((68, 2), np.float32)])
    mass = np.concatenate([mass, np.zeros(68, np.float32)])
    tree, nn, nin, pos_k, mass_k = O.bh_build(pos, mass)
    assert nin == 768 - 4  # four killed, (-500,-500) stays inside
    assert np.all(pos_k[:4] == -1000) and np.all(mass_k[:4] == 0)
    assert np.array_equal(pos_k[4], [-500, -500])
    _compare_tree(tree, H.brute_tree(pos_k, mass_k))
    # dyadic: the root COM equals the exact (float64) centre of mass
    m = mass_k.astype(np.float64)
    com = (pos_k.astype(np.float64) * m[:, None]).sum(0) / m.sum()
    assert np.allclose(tree["node_pos"][0], com, rtol=0, atol=1e-6)
    assert tree["node_mass"][0] == np.float32(m.sum())


def test_bh_build_degenerate_inputs():
    # one particle: the root is the particle node itself
    pos = np.zeros((64, 2), np.float32); pos[1:] = 900  # 63 killed
    mass = np.zeros(64, np.float32); mass[0] = 3
    tree, nn, nin, _, _ = O.bh_build(pos, mass)
    assert (nn, nin) == (1, 1) and tree["counts"][0] == 1
    # nothing inside: no nodes
    tree, nn, nin, _, _ = O.bh_build(np.full((64, 2), 900, np.float32), np.ones(64, np.float32))
    assert (nn, nin) == (0, 0)
    assert np.all(tree["sorted_src"] == 64)
    # all in one leaf cell: root chain of 10 single-child cells, then the leaf
    pos = np.full((64, 2), 0.1, np.float32)
    tree, nn, nin, _, _ = O.bh_build(pos, np.ones(64, np.float32))
    assert nn == 64 + 11
    assert list(tree["counts"][:11]) == [75 - d for d in range(11)]


# ---- Barnes-Hut traversal ------------------------------------------------------------

def _input(n=640, seed=7, spread=30.0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, spread, size=(n, 2)).astype(np.float32)
    mass = rng.uniform(1, 10, size=n).astype(np.float32)
    return pos, mass


# ---- adaptive bounding box (opt-in, beyond the reference) ---------------------

def test_bh_box_formula_and_cells():
    rng = np.random.default_rng(21)
    pos = rng.normal(0, 300, size=(3000, 2)).astype(np.float32)
    pos[:5] = [[2500, -40], [-1800, 900], [np.nan, 0], [0, np.inf], [700, 700]]
    box = O.bh_box(pos)
    assert np.array_equal(box, H.box_numpy(pos))
    fin = np.isfinite(pos).all(1)
    lo, hi = pos[fin].min(0), pos[fin].max(0)
    assert box[2] >= 0.5 * max(hi - lo)  # the square holds every finite particle
    cells = [O.cell20_box(float(x), float(y), box) for x, y in pos]
    assert cells[2] == cells[3] == 0xFFFFFFFF          # non-finite: outside
    assert all(c != 0xFFFFFFFF for c, f in zip(cells, fin) if f)
    # the extreme particles map strictly inside the 1024 x 1024 grid, without the clamp
    q = ((((pos[fin] - box[:2]) / box[2]) + np.float32(1)) * np.float32(0.5)) * np.float32(1024)
    assert q.min() >= 0 and q.max() < 1023.99
    # degenerate inputs
    assert np.array_equal(O.bh_box(np.zeros((64, 2), np.float32)), H.box_numpy(np.zeros((64, 2))))
    assert O.bh_box(np.zeros((64, 2), np.float32))[2] == np.float32(1e-3)


@pytest.mark.parametrize("seed", [22, 23])
def test_bh_build_box_matches_recursive_quadtree(seed):
    """Over the adaptive box the dense build equals the recursive quadtree of
    the same box (explicit geometry, radii r0/4^d); particles far beyond the
    reference's +-500 stay in, non-finite ones are killed."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, 400, size=(1500, 2)).astype(np.float32)
    pos[:3] = [[1900, 10], [-1200, -1500], [np.nan, 4]]
    mass = rng.uniform(1, 10, 1500).astype(np.float32)
    box = O.bh_box(pos)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass, box=box)
    assert nin == 1499 and mk[2] == 0 and mk[0] > 0 and mk[1] > 0
    _compare_tree(tree, H.brute_tree(pos, mass, box=box))
    assert tree["radii"][0] == box[3]


# ---- sparse build, 15-bit leaves (opt-in, beyond the reference) ------------------

def _sparse_inputs():
    rng = np.random.default_rng(31)
    pos = rng.normal(0, 200, size=(2000, 2)).astype(np.float32)
    pos[:600] = rng.normal(30, 0.1, size=(600, 2))         # one dense knot
    pos[600:610] = [[700, 0], [0, -900], [np.nan, 1], [500, 500], [-500, -500],
                    [3, 3], [3, 3], [3, 3], [-0.0, 0.0], [0.0, -0.0]]
    yield "knot-2000", pos, rng.uniform(1, 10, 2000).astype(np.float32)
    pos, _, mass = H.dyadic_particles(700, seed=32)
    yield "dyadic-700", pos, mass
    yield "one-point-64", np.full((64, 2), 0.1, np.float32), np.ones(64, np.float32)
    yield "single", np.array([[5.0, 6.0]], np.float32), np.array([2.0], np.float32)
    yield "empty", np.full((64, 2), 900, np.float32), np.ones(64, np.float32)


@pytest.mark.parametrize("case", range(5))
@pytest.mark.parametrize("boxed", [False, True])
def test_sparse_build_at_10_bits_is_the_dense_build(case, boxed):
    """The sparse recursion evaluates the dense pyramid's rule only where
    particles are, so at 10 bits it must reproduce the reference restatement
    byte for byte: node arrays, sorted sources, kills."""
    name, pos, mass = list(_sparse_inputs())[case]
    box = O.bh_box(pos) if boxed else None
    a = O.bh_build(pos, mass, box=box)
    b = O.bh_build(pos, mass, box=box, bits=10)
    assert a[1:3] == b[1:3], name
    for k in ("counts", "radii", "node_pos", "node_mass"):
        assert a[0][k].tobytes() == b[0][k].tobytes(), k
    assert np.array_equal(a[0]["sorted_src"], b[0]["sorted_src"])
    assert a[3].tobytes() == b[3].tobytes() and a[4].tobytes() == b[4].tobytes()


@pytest.mark.parametrize("case", range(5))
@pytest.mark.parametrize("boxed", [False, True])
def test_sparse_build_15_bits_matches_recursive_quadtree(case, boxed):
    name, pos, mass = list(_sparse_inputs())[case]
    box = O.bh_box(pos) if boxed else None
    tree, nn, nin, pk, mk = O.bh_build(pos, mass, box=box, bits=15)
    _compare_tree(tree, H.brute_tree(pos, mass, box=box, bits=15))
    if nn:
        r0 = box[3] if boxed else np.float32(500000.0)
        assert tree["radii"][0] == (r0 if nn > 1 else 0)


def test_cell_bits_prefix_and_bitloop():
    """15-bit codes: the bit-loop Morton of the truncated coordinate, and
    code >> 10 is the 20-bit cell (fixed domain and adaptive box)."""
    rng = np.random.default_rng(33)
    pos = rng.uniform(-520, 520, size=(4000, 2)).astype(np.float32)
    pos[:6] = [[-500, -500], [499.99997, 0], [500, 0], [0, -500.00003], [1e-30, -1e-30], [0, 0]]
    q = (((pos / H.SIM_SIZE) + np.float32(1)) * np.float32(0.5)) * np.float32(32768)
    ok = (q >= 0).all(1) & (q < 32768).all(1)
    want = np.where(ok, H.morton_bitloop(np.where(ok, q[:, 0], 0).astype(np.uint32),
                                         np.where(ok, q[:, 1], 0).astype(np.uint32), bits=15),
                    np.uint32(0xFFFFFFFF))
    got = np.array([O.cell_bits(float(x), float(y)) for x, y in pos], np.uint32)
    assert np.array_equal(got, want)
    c20 = np.array([O.cell20(float(x), float(y)) for x, y in pos], np.uint32)
    assert np.array_equal(np.where(got == 0xFFFFFFFF, got, got >> 10), c20)
    box = O.bh_box(pos)
    gb = np.array([O.cell_bits(float(x), float(y), box) for x, y in pos], np.uint32)
    cb = np.array([O.cell20_box(float(x), float(y), box) for x, y in pos], np.uint32)
    assert np.array_equal(gb >> 10, cb)


def test_sparse_15_bit_invariants():
    """counts[0] = N_in + #cells of >= 2 particles over depths 0..15 (from the
    codes alone); the knot that is one depth-10 bucket of 600 particles is
    split into far smaller leaves."""
    name, pos, mass = next(_sparse_inputs())
    tree, nn, nin, pk, _ = O.bh_build(pos, mass, bits=15)
    codes = np.array([O.cell_bits(float(x), float(y)) for x, y in pos], np.uint32)
    codes = codes[codes != 0xFFFFFFFF]
    ncell = 0
    for d in range(16):
        _, cnt = np.unique(codes >> np.uint32(2 * (15 - d)), return_counts=True)
        ncell += int((cnt >= 2).sum())
    assert nin == codes.size and tree["counts"][0] == nin + ncell == nn
    _, b10 = np.unique(codes >> np.uint32(10), return_counts=True)
    _, b15 = np.unique(codes, return_counts=True)
    assert b10.max() >= 300 and b15.max() <= b10.max() // 4
    assert tree["radii"][tree["radii"] > 0].min() == np.float32(500000.0) / np.float32(4.0**15)


@pytest.mark.parametrize("W", [1, 32, 64])
@pytest.mark.parametrize("theta", [0.5, 1.0])
def test_bh_traverse_matches_recursive_group_vote(W, theta):
    pos, mass = _input(n=320 if W == 1 else 640)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    eps2, t2 = H.sq32(0.2), H.sq32(theta)
    acc, V, A, _, _ = O.bh_traverse(tree, pk, mk, eps2, t2, W=W)
    bacc, bV, bA = H.brute_group_forces(tree, pk, eps2, t2, W)
    assert (V, A) == (bV, bA)
    assert np.array_equal(acc, bacc)


def test_bh_theta_zero_is_direct_sum_in_sorted_order():
    pos, mass = _input(n=512, seed=8)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    eps2 = H.sq32(0.2)
    acc, V, A, _, _ = O.bh_traverse(tree, pk, mk, eps2, 0.0, W=64)
    order = tree["sorted_src"][tree["sorted_src"] < len(mk)]
    # direct sum over sources in sorted order == particle nodes in pre-order;
    # ds[k] is the acceleration of particle order[k]
    ds = O.ds_accel_f32(pk[order], mk[order], eps2)
    assert np.array_equal(acc[order], ds)
    assert A == nin * nin  # every real lane accepts every particle node


def test_bh_theta_infinite_is_pull_of_global_com():
    pos, mass = _input(n=256, seed=9)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    eps2 = H.sq32(0.2)
    acc, V, A, _, _ = O.bh_traverse(tree, pk, mk, eps2, 1e30, W=64)
    m = mk.astype(np.float64)
    com = (pk.astype(np.float64) * m[:, None]).sum(0) / m.sum()
    d = com[None, :] - pk.astype(np.float64)
    r2 = (d ** 2).sum(1) + eps2
    ref = d * (m.sum() / r2 ** 1.5)[:, None]
    assert H.system_rel_err(acc, ref) < 1e-5
    assert V == A == nin  # one visit per lane: the root


def test_bh_traverse_f64_accumulation_is_exact_on_the_same_terms():
    """The double-precision sum of the walk's accepted terms (the yardstick
    for GPU walks that add the same terms in another order): at theta = 0 it
    is the FP64 direct sum; at theta -> inf it is the exact pull of the root
    node the tree built; in between it stays within FP32 rounding of the FP32
    walk."""
    pos, mass = _input(n=512, seed=12)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    eps2 = H.sq32(0.2)
    inside = tree["sorted_src"][tree["sorted_src"] < len(mk)]
    acc, V, A, _, _, a64 = O.bh_traverse(tree, pk, mk, eps2, 0.0, W=64, f64=True)
    ds64 = O.ds_accel_f64(pk[inside], mk[inside], eps2)
    assert np.allclose(a64[inside], ds64, rtol=1e-12, atol=1e-12)
    acc, V, A, _, _, a64 = O.bh_traverse(tree, pk, mk, eps2, 1e30, W=64, f64=True)
    root = tree["node_pos"][0].astype(np.float64)
    d = root[None, :] - pk[inside].astype(np.float64)
    q = 1.0 / np.sqrt((d ** 2).sum(1) + np.float64(eps2))
    ref = d * (np.float64(tree["node_mass"][0]) * q ** 3)[:, None]
    assert np.allclose(a64[inside], ref, rtol=1e-14, atol=0)
    acc, V, A, _, _, a64 = O.bh_traverse(tree, pk, mk, eps2, 1.0, W=64, f64=True)
    assert H.system_rel_err(acc[inside], a64[inside]) < 1e-6
    assert H.per_particle_rel(acc[inside], a64[inside]).max() < 1e-5


def test_bh_error_against_f64_shrinks_with_theta():
    pos, mass = _input(n=1024, seed=10, spread=20.0)
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    eps2 = H.sq32(0.2)
    a64 = O.ds_accel_f64(pk, mk, eps2)
    errs = []
    for th in (1.0, 0.5, 0.25):
        acc, *_ = O.bh_traverse(tree, pk, mk, eps2, H.sq32(th), W=64)
        errs.append(H.system_rel_err(acc, a64))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 5e-3


def test_bh_step_integrates_and_keeps_killed_velocity():
    pos, mass = _input(n=200, seed=11)
    pos[0] = [800, 0]
    pos = np.vstack([pos, np.zeros((56, 2), np.float32)])
    mass = np.concatenate([mass, np.zeros(56, np.float32)])
    vel = np.ones_like(pos)
    po, vo, mo, V, A = O.bh_step(pos, vel, mass, 1.0, 1e-3, H.sq32(0.2), 1.0, 64)
    assert np.array_equal(po[0], [-1000, -1000]) and mo[0] == 0
    assert np.array_equal(vo[0], vel[0])
    # integrate: v' = v + a*(G*dt); p' = p + (v+v')*(0.5*dt), per ForceShader.comp:145-146
    tree, nn, nin, pk, mk = O.bh_build(pos, mass)
    acc, *_ = O.bh_traverse(tree, pk, mk, H.sq32(0.2), 1.0, W=64)
    kv = np.float32(1.0) * np.float32(1e-3)
    kp = np.float32(0.5) * np.float32(1e-3)
    nv = (vel[1:] + acc[1:] * kv).astype(np.float32)
    np_ = (pk[1:] + ((vel[1:] + nv).astype(np.float32) * kp)).astype(np.float32)
    assert np.array_equal(vo[1:], nv)
    assert np.array_equal(po[1:], np_)


@pytest.mark.parametrize("G,dt", [(1.0, 1e-3), (0.6, 4e-3)])
def test_bh_step_theta0_tracks_f64_direct_sum_step(G, dt):
    """At theta = 0 every cell opens, so one Barnes-Hut step is a direct-sum
    step in another summation order: it must agree with the FP64 yardstick
    (pinned above against Newton's law) to FP32 rounding, which pins the
    integrate's G, dt and half-step factors independently of its code."""
    pos, mass = _input(n=192, seed=13, spread=10.0)
    rng = np.random.default_rng(14)
    vel = rng.normal(0, 2, size=pos.shape).astype(np.float32)
    po, vo, mo, V, A = O.bh_step(pos, vel, mass, G, dt, H.sq32(0.2), 0.0, 64)
    p64, v64 = O.ds_steps_f64(pos, vel, mass, G, dt, H.sq32(0.2), 1)
    dv = np.abs(v64 - vel).max()
    assert dv > 1e-2  # the step does feel the forces
    assert np.abs(vo - v64).max() < 1e-5 * dv + 4 * np.finfo(np.float32).eps * np.abs(vel).max()
    assert np.abs(po - p64).max() < 4 * np.finfo(np.float32).eps * np.abs(pos).max()
    assert np.array_equal(mo, mass)
