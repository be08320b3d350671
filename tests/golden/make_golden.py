"""Writes kat.json, the known-answer vectors the oracle is checked against
(tests/test_golden.py).  The reference (C++/GLSL over Vulkan) ships no test
vectors or fixtures, so these come from closed forms and from brute force that
shares no code with the oracle: numpy bit loops for the cells, analytic
two-body and symmetric-ring forces in float64, and the recursive pointer
quadtree of tests/helpers.py on dyadic inputs (every FP32 sum exact).  FP32
values are stored as their uint32 bit patterns so NaN and -0 survive.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import helpers as H  # noqa: E402


def bits(a) -> list:
    return np.asarray(a, np.float32).view(np.uint32).ravel().tolist()


def cells() -> dict:
    """Cells of edge positions: both domain edges, the FP32 tie just below +500
    (q rounds to 1024: outside), signed zero, NaN, huge and tiny values."""
    f = np.float32
    below = np.nextafter(f(500.0), f(0.0))
    pts = [(-500.0, -500.0), (500.0, 0.0), (below, 0.0), (0.0, below), (0.0, 0.0), (-0.0, 0.0),
           (499.02, -499.9), (np.nan, 1.0), (1.0, np.nan), (1e30, 0.0), (-500.0001, 3.0),
           (12.5, -7.25), (250.0, 250.0), (-250.0, 125.5), (0.48828125, 0.48828125),
           (0.48828124, -0.48828125), (3e-39, -3e-39)]
    pos = np.array(pts, np.float32)
    return {"pos_bits": bits(pos), "cell": H.cells_numpy(pos).tolist()}


def two_body() -> dict:
    """Masses m0, m1 at x = 0 and x = d: a0 = m1 d / (d^2 + eps^2)^(3/2) along +x,
    a1 = -m0 d / (d^2 + eps^2)^(3/2), exactly, with eps^2 the FP32 square."""
    eps2 = float(H.sq32(0.2))
    cases = []
    for m0, m1, d in ((1.0, 1.0, 1.0), (3.0, 7.0, 0.25), (10.0, 1.0, 40.0), (2.0, 5.0, 0.01)):
        k = d / (d * d + eps2) ** 1.5
        cases.append({"pos_bits": bits([[0.0, 0.0], [d, 0.0]]), "mass_bits": bits([m0, m1]),
                      "acc": [[m1 * k, 0.0], [-m0 * k, 0.0]]})
    return {"eps2_bits": bits([eps2])[0], "cases": cases}


def ring() -> dict:
    """A particle at the centre of a ring of 8 equal masses at 90-degree steps of
    a square and its diagonals: the centre's acceleration is 0 by symmetry; a
    ring particle's is the analytic pull of the other 8."""
    eps2 = float(H.sq32(0.2))
    pts = [(0.0, 0.0), (4.0, 0.0), (-4.0, 0.0), (0.0, 4.0), (0.0, -4.0),
           (4.0, 4.0), (-4.0, -4.0), (4.0, -4.0), (-4.0, 4.0)]
    mass = [5.0] + [2.0] * 8
    p = np.array(pts, np.float64)
    a1 = np.zeros(2)
    for j in range(9):
        if j == 1:
            continue
        d = p[j] - p[1]
        a1 += mass[j] * d / (d @ d + eps2) ** 1.5
    return {"eps2_bits": bits([eps2])[0], "pos_bits": bits(pts), "mass_bits": bits(mass),
            "acc0": [0.0, 0.0], "acc1": a1.tolist()}


def tree() -> dict:
    """Node arrays of a dyadic particle set (with two particles outside the
    domain) by the recursive pointer quadtree: exact in any summation order."""
    pos, _, mass = H.dyadic_particles(40, seed=901, spread=60)
    pos[3] = (600.0, 0.0)
    pos[17] = (0.0, -512.0)
    pos[20:24] = pos[19]  # a shared leaf
    t = H.brute_tree(pos, mass)
    out = {"pos_bits": bits(pos), "mass_bits": bits(mass)}
    for k in ("counts", "sorted_src"):
        out[k] = np.asarray(t[k]).astype(np.int64).tolist()
    for k in ("radii", "node_mass", "node_pos"):
        out[k + "_bits"] = bits(t[k])
    return out


def main() -> None:
    kat = {"cells": cells(), "two_body": two_body(), "ring": ring(), "tree": tree()}
    path = os.path.join(HERE, "kat.json")
    with open(path, "w") as fh:
        json.dump(kat, fh, indent=1)
        fh.write("\n")
    print(path)


if __name__ == "__main__":
    main()
