"""ctypes front of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg may import this
module; it is the checker, never the product.  Functions mirror oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
CAPACITY_EXTRA = 1398101  # BarnesHutSimulation.cpp:57

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        u64 = C.POINTER(C.c_uint64)
        L.oracle_ds_accel_f32.argtypes = [C.c_size_t, fp, fp, C.c_float, fp, C.c_int]
        L.oracle_ds_accel_f64.argtypes = [C.c_size_t, fp, fp, C.c_double, dp, C.c_int]
        L.oracle_ds_accel_f64_at.argtypes = [C.c_size_t, fp, fp, C.c_double, up, C.c_size_t,
                                             dp, C.c_int]
        L.oracle_ds_step.argtypes = [C.c_size_t, fp, fp, fp, fp, fp, C.c_float, C.c_float,
                                     C.c_float, C.c_int]
        L.oracle_ds_step_f64.argtypes = [C.c_size_t, dp, dp, dp, C.c_double, C.c_double,
                                         C.c_double, C.c_int]
        L.oracle_cell20.argtypes = [C.c_float, C.c_float]
        L.oracle_cell20.restype = C.c_uint32
        L.oracle_bh_capacity.argtypes = [C.c_size_t]
        L.oracle_bh_capacity.restype = C.c_size_t
        L.oracle_bh_box.argtypes = [C.c_size_t, fp, fp]
        L.oracle_cell20_box.argtypes = [C.c_float, C.c_float, fp]
        L.oracle_cell20_box.restype = C.c_uint32
        L.oracle_bh_build_box.argtypes = [C.c_size_t, fp, fp, fp, fp, fp, C.c_void_p,
                                          C.POINTER(C.c_size_t)]
        L.oracle_bh_build_box.restype = C.c_long
        L.oracle_bh_build.argtypes = [C.c_size_t, fp, fp, fp, fp, C.c_void_p,
                                      C.POINTER(C.c_size_t)]
        L.oracle_bh_build.restype = C.c_long
        L.oracle_cell_bits.argtypes = [C.c_float, C.c_float, fp, C.c_uint]
        L.oracle_cell_bits.restype = C.c_uint32
        L.oracle_bh_capacity_sparse.argtypes = [C.c_size_t, C.c_uint]
        L.oracle_bh_capacity_sparse.restype = C.c_size_t
        L.oracle_bh_build_sparse.argtypes = [C.c_size_t, fp, fp, fp, fp, fp, C.c_uint,
                                             C.c_void_p, C.POINTER(C.c_size_t)]
        L.oracle_bh_build_sparse.restype = C.c_long
        L.oracle_bh_traverse.argtypes = [C.c_size_t, C.c_uint32, fp, fp, C.c_void_p,
                                         C.c_size_t, C.c_float, C.c_float, fp, dp, u64, u64,
                                         u64, C.POINTER(C.c_uint8), C.c_int]
        L.oracle_bh_integrate.argtypes = [C.c_size_t, fp, fp, fp, fp, fp, fp, fp, C.c_float,
                                          C.c_float]
        L.oracle_bh_step.argtypes = [C.c_size_t, fp, fp, fp, fp, fp, fp, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_uint32, u64, u64, C.c_int]
        L.oracle_bh_step.restype = C.c_int
        _lib = L
    return _lib


class _Tree(C.Structure):
    _fields_ = [("counts", C.POINTER(C.c_uint32)), ("radii", C.POINTER(C.c_float)),
                ("node_pos", C.POINTER(C.c_float)), ("node_mass", C.POINTER(C.c_float)),
                ("sorted_src", C.POINTER(C.c_uint32)), ("cap", C.c_size_t)]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nthreads_default() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def ds_accel_f32(pos, mass, eps2: float, nthreads: int = 0) -> np.ndarray:
    pos, mass = _f32(pos), _f32(mass)
    n = mass.shape[0]
    acc = np.zeros((n, 2), np.float32)
    lib().oracle_ds_accel_f32(n, _f(pos), _f(mass), eps2, _f(acc), nthreads or nthreads_default())
    return acc


def ds_accel_f64(pos, mass, eps2: float, nthreads: int = 0) -> np.ndarray:
    pos, mass = _f32(pos), _f32(mass)
    n = mass.shape[0]
    acc = np.zeros((n, 2), np.float64)
    lib().oracle_ds_accel_f64(n, _f(pos), _f(mass), float(eps2),
                              acc.ctypes.data_as(C.POINTER(C.c_double)),
                              nthreads or nthreads_default())
    return acc


def ds_accel_f64_at(pos, mass, eps2: float, targets, nthreads: int = 0) -> np.ndarray:
    """FP64 direct-sum accelerations of the listed targets only, (k, 2)."""
    pos, mass = _f32(pos), _f32(mass)
    t = np.ascontiguousarray(targets, np.uint32)
    acc = np.zeros((t.shape[0], 2), np.float64)
    lib().oracle_ds_accel_f64_at(mass.shape[0], _f(pos), _f(mass), float(eps2),
                                 t.ctypes.data_as(C.POINTER(C.c_uint32)), t.shape[0],
                                 acc.ctypes.data_as(C.POINTER(C.c_double)),
                                 nthreads or nthreads_default())
    return acc


def ds_step(pos, vel, mass, G: float, dt: float, eps2: float, nthreads: int = 0):
    pos, vel, mass = _f32(pos), _f32(vel), _f32(mass)
    n = mass.shape[0]
    po = np.zeros_like(pos)
    vo = np.zeros_like(vel)
    lib().oracle_ds_step(n, _f(pos), _f(vel), _f(mass), _f(po), _f(vo), G, dt, eps2,
                         nthreads or nthreads_default())
    return po, vo


def ds_steps_f64(pos, vel, mass, G: float, dt: float, eps2: float, steps: int,
                 nthreads: int = 0):
    """`steps` direct-sum steps in FP64; returns (pos, vel) as float64."""
    p = np.ascontiguousarray(pos, np.float64).copy()
    v = np.ascontiguousarray(vel, np.float64).copy()
    m = np.ascontiguousarray(mass, np.float64)
    n = m.shape[0]
    dp = C.POINTER(C.c_double)
    for _ in range(steps):
        lib().oracle_ds_step_f64(n, p.ctypes.data_as(dp), v.ctypes.data_as(dp),
                                 m.ctypes.data_as(dp), G, dt, eps2,
                                 nthreads or nthreads_default())
    return p, v


def cell20(x: float, y: float) -> int:
    return int(lib().oracle_cell20(x, y))


def bh_box(pos) -> np.ndarray:
    """Adaptive bounding box {cx, cy, h, r0} of the finite positions."""
    pos = _f32(pos)
    box = np.zeros(4, np.float32)
    lib().oracle_bh_box(pos.shape[0], _f(pos), _f(box))
    return box


def cell20_box(x: float, y: float, box) -> int:
    box = _f32(box)
    return int(lib().oracle_cell20_box(x, y, _f(box)))


def cell_bits(x: float, y: float, box=None, bits: int = 15) -> int:
    """Cell on a 2^bits grid per axis (fixed domain, or the box {cx, cy, h, r0})."""
    if box is None:
        return int(lib().oracle_cell_bits(x, y, None, bits))
    bx = _f32(box)
    return int(lib().oracle_cell_bits(x, y, _f(bx), bits))


def bh_build(pos, mass, box=None, bits=None):
    """Returns (tree dict, n_nodes, n_inside, pos_after_kill, mass_after_kill);
    with box ({cx, cy, h, r0}) the build is over that box instead of the fixed
    [-500, 500) domain.  bits=None is the dense pyramid restatement of the
    reference; bits=10 or 15 the sparse build with leaves at that depth."""
    pos, mass = _f32(pos).copy(), _f32(mass).copy()
    na = mass.shape[0]
    if bits is None:
        cap = int(lib().oracle_bh_capacity(na))
    else:
        cap = int(lib().oracle_bh_capacity_sparse(na, bits))
    counts = np.zeros(cap, np.uint32)
    radii = np.zeros(cap, np.float32)
    npos = np.zeros((cap, 2), np.float32)
    nmass = np.zeros(cap, np.float32)
    src = np.zeros(na, np.uint32)
    t = _Tree(counts.ctypes.data_as(C.POINTER(C.c_uint32)), _f(radii), _f(npos), _f(nmass),
              src.ctypes.data_as(C.POINTER(C.c_uint32)), cap)
    inside = C.c_size_t(0)
    if bits is not None:
        bx = None if box is None else _f32(box)
        nn = lib().oracle_bh_build_sparse(na, _f(pos), _f(mass), None, None,
                                          None if bx is None else _f(bx), bits, C.byref(t),
                                          C.byref(inside))
    elif box is None:
        nn = lib().oracle_bh_build(na, _f(pos), _f(mass), None, None, C.byref(t), C.byref(inside))
    else:
        bx = _f32(box)
        nn = lib().oracle_bh_build_box(na, _f(pos), _f(mass), None, None, _f(bx), C.byref(t),
                                       C.byref(inside))
    if nn < 0:
        raise RuntimeError("oracle_bh_build: capacity exceeded")
    tree = {"counts": counts[:nn].copy(), "radii": radii[:nn].copy(),
            "node_pos": npos[:nn].copy(), "node_mass": nmass[:nn].copy(),
            "sorted_src": src, "total": int(nn)}
    return tree, int(nn), int(inside.value), pos, mass


def bh_traverse(tree: dict, pos, mass, eps2: float, theta2: float, W: int = 64,
                nthreads: int = 0, f64: bool = False):
    """Returns (acc[na,2], visits, accepts, ambiguous_groups, group_flags), and
    with f64=True also the same accepted terms summed in double, acc64[na,2]."""
    pos, mass = _f32(pos), _f32(mass)
    na = mass.shape[0]
    nn = int(tree["total"])
    counts = np.ascontiguousarray(tree["counts"], np.uint32)
    radii = _f32(tree["radii"])
    npos = _f32(tree["node_pos"]).reshape(-1, 2)
    nmass = _f32(tree["node_mass"])
    src = np.ascontiguousarray(tree["sorted_src"], np.uint32)
    t = _Tree(counts.ctypes.data_as(C.POINTER(C.c_uint32)), _f(radii), _f(npos), _f(nmass),
              src.ctypes.data_as(C.POINTER(C.c_uint32)), max(nn, 1))
    acc = np.zeros((na, 2), np.float32)
    acc64 = np.zeros((na, 2), np.float64) if f64 else None
    v = C.c_uint64(0)
    a = C.c_uint64(0)
    am = C.c_uint64(0)
    flags = np.zeros(max(na // W, 1), np.uint8)
    lib().oracle_bh_traverse(na, W, _f(pos), _f(mass), C.byref(t), nn, eps2, theta2, _f(acc),
                             acc64.ctypes.data_as(C.POINTER(C.c_double)) if f64 else None,
                             C.byref(v), C.byref(a), C.byref(am),
                             flags.ctypes.data_as(C.POINTER(C.c_uint8)),
                             nthreads or nthreads_default())
    out = (acc, int(v.value), int(a.value), int(am.value), flags)
    return out + (acc64,) if f64 else out


def bh_step(pos, vel, mass, G: float, dt: float, eps2: float, theta2: float, W: int = 64,
            nthreads: int = 0):
    """One Barnes-Hut step.  Returns (pos_out, vel_out, mass_out, visits, accepts)."""
    pos, vel, mass = _f32(pos).copy(), _f32(vel).copy(), _f32(mass).copy()
    na = mass.shape[0]
    po = np.zeros_like(pos)
    vo = np.zeros_like(vel)
    mo = mass.copy()
    v = C.c_uint64(0)
    a = C.c_uint64(0)
    rc = lib().oracle_bh_step(na, _f(pos), _f(vel), _f(mass), _f(po), _f(vo), _f(mo), G, dt,
                              eps2, theta2, W, C.byref(v), C.byref(a),
                              nthreads or nthreads_default())
    if rc != 0:
        raise RuntimeError(f"oracle_bh_step failed: {rc}")
    return po, vo, mo, int(v.value), int(a.value)


def f32(x: float) -> float:
    return float(np.float32(x))


def square_f32(x: float) -> float:
    """x*x rounded to FP32, as the reference squares softening and theta on the host."""
    v = np.float32(x)
    return float(v * v)
