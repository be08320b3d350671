"""Shared test helpers: brute-force restatements that pin the oracle, tolerance
checks, and seeded inputs."""
from __future__ import annotations

import numpy as np

SIM_SIZE = np.float32(500.0)
TREE = 1024


def f32(x) -> np.float32:
    return np.float32(x)


def sq32(x) -> float:
    v = np.float32(x)
    return float(v * v)


# ---- Morton by bit loop (independent of the magic-mask form) ----------------

def morton_bitloop(ix: np.ndarray, iy: np.ndarray, bits: int = 10) -> np.ndarray:
    ix = ix.astype(np.uint64)
    iy = iy.astype(np.uint64)
    out = np.zeros(ix.shape, np.uint64)
    for b in range(bits):
        out |= ((ix >> b) & 1) << (2 * b)
        out |= ((iy >> b) & 1) << (2 * b + 1)
    return out.astype(np.uint32)


def cells_numpy(pos: np.ndarray) -> np.ndarray:
    """20-bit cell per particle in FP32 numpy arithmetic (0xffffffff = out)."""
    p = pos.astype(np.float32)
    q = ((p / SIM_SIZE) + np.float32(1.0)) * np.float32(0.5)
    q = q * np.float32(TREE)
    with np.errstate(invalid="ignore"):
        inside = (q[:, 0] >= 0) & (q[:, 0] < TREE) & (q[:, 1] >= 0) & (q[:, 1] < TREE)
    qi = np.where(inside[:, None], q, 0).astype(np.uint32)
    c = morton_bitloop(qi[:, 0], qi[:, 1])
    return np.where(inside, c, np.uint32(0xFFFFFFFF)).astype(np.uint32)


# ---- recursive pointer quadtree (brute force for small N) --------------------

class _Cell:
    __slots__ = ("depth", "members", "children")

    def __init__(self, depth, members):
        self.depth = depth
        self.members = members  # particle indices, ascending
        self.children = []


def box_numpy(pos: np.ndarray) -> np.ndarray:
    """The adaptive box {cx, cy, h, r0} restated in numpy FP32: centre of the
    finite extremes, half-size = max extent x (0.5 + 2^-9), at least 1e-3."""
    p = pos.astype(np.float32)
    fin = np.isfinite(p).all(axis=1)
    if not fin.any():
        lo = hi = np.zeros(2, np.float32)
    else:
        lo, hi = p[fin].min(axis=0), p[fin].max(axis=0)
    c = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    w = np.float32(max(np.float32(hi[0] - lo[0]), np.float32(hi[1] - lo[1])))
    h = np.float32(w * np.float32(0.501953125))
    if not h >= np.float32(1e-3):
        h = np.float32(1e-3)
    return np.array([c[0], c[1], h, np.float32(np.float32(h * h) * np.float32(2.0))], np.float32)


def brute_tree(pos: np.ndarray, mass: np.ndarray, box=None, bits: int = 10):
    """Builds the reference topology by recursive subdivision of the domain
    (explicit geometry, no Morton codes): quadrant by comparing the normalised
    coordinate against the cell's midpoint.  Returns node arrays in pre-order
    and the sorted source list, with FP32 sums in the reference association.
    bits = 15 puts the leaves on a 32768^2 grid (depth 15) instead of 1024^2."""
    na = len(mass)
    p = pos.astype(np.float32)
    TREE = 1 << bits
    if box is None:
        q = (((p / SIM_SIZE) + np.float32(1.0)) * np.float32(0.5)) * np.float32(TREE)
        with np.errstate(invalid="ignore"):
            ok = (q[:, 0] >= 0) & (q[:, 0] < TREE) & (q[:, 1] >= 0) & (q[:, 1] < TREE)
        inside = [i for i in range(na) if ok[i]]
        # integer leaf coordinates from the same normalisation
        ij = {i: (int(q[i, 0]), int(q[i, 1])) for i in inside}
        r0 = np.float32(500000.0)
    else:
        b = np.asarray(box, np.float32)
        inside = [i for i in range(na) if np.isfinite(p[i]).all()]
        with np.errstate(invalid="ignore"):
            q = ((((p - b[:2]) / b[2]) + np.float32(1.0)) * np.float32(0.5)) * np.float32(TREE)
        ij = {i: (min(int(max(q[i, 0], 0)), TREE - 1), min(int(max(q[i, 1], 0)), TREE - 1))
              for i in inside}
        r0 = b[3]

    def split(members, depth, x0, y0):
        half = TREE >> (depth + 1)
        quads = [[], [], [], []]
        for i in members:
            ix, iy = ij[i]
            k = (1 if ix >= x0 + half else 0) | (2 if iy >= y0 + half else 0)
            quads[k].append(i)
        origins = [(x0, y0), (x0 + half, y0), (x0, y0 + half), (x0 + half, y0 + half)]
        return quads, origins

    counts, radii, npos, nmass, order = [], [], [], [], []

    def leaf_sums(members):
        sx = np.float32(0); sy = np.float32(0); sm = np.float32(0)
        for i in members:  # ascending index
            sx = np.float32(sx + np.float32(p[i, 0] * mass[i]))
            sy = np.float32(sy + np.float32(p[i, 1] * mass[i]))
            sm = np.float32(sm + np.float32(mass[i]))
        return sx, sy, sm

    def sums(members, depth, x0, y0):
        if depth == bits:
            return leaf_sums(members)
        quads, origins = split(members, depth, x0, y0)
        acc = []
        for k in range(4):
            if quads[k]:
                acc.append(sums(quads[k], depth + 1, *origins[k]))
            else:
                acc.append((np.float32(0), np.float32(0), np.float32(0)))
        out = []
        for c in range(3):
            v = np.float32(np.float32(acc[0][c] + acc[1][c]) + acc[2][c])
            out.append(np.float32(v + acc[3][c]))
        return tuple(out)

    def emit(members, depth, x0, y0):
        if len(members) == 1:
            i = members[0]
            order.append(i)
            counts.append(1); radii.append(np.float32(0))
            npos.append((p[i, 0], p[i, 1])); nmass.append(mass[i])
            return 1
        slot = len(counts)
        sx, sy, sm = sums(members, depth, x0, y0)
        cx, cy = (sx / sm, sy / sm) if sm != 0 else (sx, sy)
        counts.append(0)
        radii.append(np.float32(np.float32(r0) / np.float32(4 ** depth)))
        npos.append((np.float32(cx), np.float32(cy))); nmass.append(sm)
        size = 1
        if depth == bits:
            for i in members:
                order.append(i)
                counts.append(1); radii.append(np.float32(0))
                npos.append((p[i, 0], p[i, 1])); nmass.append(mass[i])
                size += 1
        else:
            quads, origins = split(members, depth, x0, y0)
            for k in range(4):
                if quads[k]:
                    size += emit(quads[k], depth + 1, *origins[k])
        counts[slot] = size
        return size

    if inside:
        emit(inside, 0, 0, 0)
    src = np.full(na, na, np.uint32)
    src[: len(order)] = order
    return {
        "counts": np.array(counts, np.uint32),
        "radii": np.array(radii, np.float32),
        "node_pos": np.array(npos, np.float32).reshape(-1, 2),
        "node_mass": np.array(nmass, np.float32),
        "sorted_src": src,
        "total": len(counts),
    }


def brute_group_forces(tree: dict, pos, eps2: float, theta2: float, W: int):
    """Recursive restatement of the group traversal on the pre-order arrays:
    visit(node) accepts the whole subtree if all real lanes accept, else visits
    the node's children (the nodes of its subtree reached by skipping)."""
    na = pos.shape[0]
    src = tree["sorted_src"]
    counts = tree["counts"]
    radii = tree["radii"]
    npos = tree["node_pos"]
    nmass = tree["node_mass"]
    total = tree["total"]
    acc = np.zeros((na, 2), np.float64)
    V = A = 0
    e2 = np.float32(eps2)
    t2 = np.float32(theta2)
    for g in range(na // W):
        lanes = src[g * W:(g + 1) * W]
        real = lanes < na
        lp = np.where(real[:, None], pos[np.minimum(lanes, na - 1)], 0).astype(np.float32)
        ax = np.zeros(W, np.float32)
        ay = np.zeros(W, np.float32)
        nreal = int(real.sum())

        def visit(u):
            nonlocal V, A
            d = (npos[u] - lp).astype(np.float32)
            r2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32) + e2).astype(np.float32)
            ok = (radii[u] <= (r2 * t2).astype(np.float32)) | ~real
            V += nreal
            if ok.all():
                A += nreal
                inv = (np.float32(1) / np.sqrt(r2)).astype(np.float32)
                s = (((nmass[u] * inv).astype(np.float32) * inv).astype(np.float32) * inv).astype(np.float32)
                ax[:] = (ax + (d[:, 0] * s).astype(np.float32)).astype(np.float32)
                ay[:] = (ay + (d[:, 1] * s).astype(np.float32)).astype(np.float32)
                return
            child = u + 1
            end = u + counts[u]
            while child < end:
                visit(child)
                child += counts[child]

        if total > 0:
            # the root's siblings: walk top-level nodes (one root unless N_in == 1)
            u = 0
            while u < total:
                visit(u)
                u += counts[u]
        for l in range(W):
            if real[l]:
                acc[lanes[l]] = (ax[l], ay[l])
    return acc.astype(np.float32), V, A


# ---- tolerance checks ----------------------------------------------------------

def system_rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def per_particle_rel(a, b) -> np.ndarray:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mag = np.linalg.norm(b, axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if len(mag) else 1.0
    return np.linalg.norm(a - b, axis=1) / np.maximum(mag, rms)


def dyadic_particles(n: int, seed: int, spread: int = 40):
    """Positions on a 2^-8 grid and small-integer masses: every FP32 sum of
    m*x is exact, so centre-of-mass sums agree bit-for-bit in any order."""
    rng = np.random.default_rng(seed)
    pos = (rng.integers(-spread * 256, spread * 256, size=(n, 2)) / 256.0).astype(np.float32)
    mass = rng.integers(1, 8, size=n).astype(np.float32)
    vel = np.zeros((n, 2), np.float32)
    return pos, vel, mass
