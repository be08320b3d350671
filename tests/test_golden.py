"""The oracle against the committed known-answer vectors (tests/golden/kat.json,
written by tests/golden/make_golden.py from closed forms and brute force that
share no code with the oracle).  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O

KAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat.json")


@pytest.fixture(scope="module")
def kat():
    with open(KAT) as fh:
        return json.load(fh)


def f32(b) -> np.ndarray:
    return np.asarray(b, np.uint32).view(np.float32)


def test_golden_cells(kat):
    """Cells of the domain edges, the FP32 tie below +500, signed zero, NaN
    and huge values (InitShader.comp:80-98, 110-117) against a bit loop."""
    pos = f32(kat["cells"]["pos_bits"]).reshape(-1, 2)
    got = [O.cell20(float(x), float(y)) for x, y in pos]
    assert got == kat["cells"]["cell"]


def test_golden_two_body(kat):
    """Two bodies: a0 = m1 d / (d^2 + eps^2)^(3/2), a1 = -m0 d / (...)
    (SimShader.comp:53-60), to the FP32 rounding of a few operations."""
    eps2 = float(f32([kat["two_body"]["eps2_bits"]])[0])
    for case in kat["two_body"]["cases"]:
        pos = f32(case["pos_bits"]).reshape(-1, 2)
        mass = f32(case["mass_bits"])
        a = O.ds_accel_f32(pos, mass, eps2).astype(np.float64)
        want = np.asarray(case["acc"])
        assert np.allclose(a, want, rtol=2e-6, atol=0), (case, a)
        assert a[0, 1] == 0.0 and a[1, 1] == 0.0


def test_golden_ring(kat):
    """The centre of a symmetric ring feels exactly zero (opposite members are
    adjacent in the sum, so their FP32 terms cancel exactly); a ring member
    feels the analytic pull of the other eight."""
    r = kat["ring"]
    eps2 = float(f32([r["eps2_bits"]])[0])
    pos = f32(r["pos_bits"]).reshape(-1, 2)
    mass = f32(r["mass_bits"])
    a = O.ds_accel_f32(pos, mass, eps2).astype(np.float64)
    assert np.array_equal(a[0], np.zeros(2))
    want = np.asarray(r["acc1"])
    assert np.linalg.norm(a[1] - want) <= 2e-6 * np.linalg.norm(want)


def test_golden_tree(kat):
    """Node arrays of a dyadic set with two particles outside the domain and a
    shared leaf: the oracle's dense build equals the recursive quadtree bit
    for bit (every FP32 sum is exact)."""
    t = kat["tree"]
    pos = f32(t["pos_bits"]).reshape(-1, 2)
    mass = f32(t["mass_bits"])
    tree, nn, _, _, _ = O.bh_build(pos, mass)
    assert nn == len(t["counts"])
    assert np.array_equal(tree["counts"], np.asarray(t["counts"], np.uint32))
    assert np.array_equal(tree["sorted_src"], np.asarray(t["sorted_src"], np.uint32))
    assert np.array_equal(tree["radii"].view(np.uint32), np.asarray(t["radii_bits"], np.uint32))
    assert np.array_equal(tree["node_mass"].view(np.uint32), np.asarray(t["node_mass_bits"], np.uint32))
    assert np.array_equal(tree["node_pos"].view(np.uint32).ravel(), np.asarray(t["node_pos_bits"], np.uint32))
