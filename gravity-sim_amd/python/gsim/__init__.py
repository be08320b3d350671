"""Python binding of libgsim (the MI355X force-and-integrate library) over its
C ABI (include/gsim.h), with the reference's object surface:

    ParticleSystem     particle set + parameters (src/Particles/ParticleSystem.hpp:58,72)
    DirectSimulation   RunSimulations / GetRequiredParticleAlignment
                       (src/Simulation/Direct/DirectSimulation.hpp:15-24)
    BarnesHutSimulation  the same for --simulation-algorithm=barnes-hut
                       (src/Simulation/BarnesHut/BarnesHutSimulation.hpp:15-24)

Errors raise GsimError carrying the library's message, as the reference throws
gsim::Exception (src/Debug/Exception.hpp:72).  There is no CPU fallback: if
libgsim.so is missing or no gfx950 device is visible, the call fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libgsim.so")

DIRECT_SUM = 0
BARNES_HUT = 1

GENERATE_RANDOM = 0
GENERATE_GALAXY = 1
GENERATE_GALAXY_COLLISION = 2
GENERATE_SYMMETRICAL_GALAXY_COLLISION = 3
GENERATE_EXPONENTIAL_DISK = 16
GENERATE_TWO_EXPONENTIAL_DISKS = 17

GENERATE_TYPES = {
    "random": GENERATE_RANDOM,
    "galaxy": GENERATE_GALAXY,
    "galaxy-collision": GENERATE_GALAXY_COLLISION,
    "symmetrical-galaxy-collision": GENERATE_SYMMETRICAL_GALAXY_COLLISION,
    "exponential-disk": GENERATE_EXPONENTIAL_DISK,
    "two-exponential-disks": GENERATE_TWO_EXPONENTIAL_DISKS,
}
ALGORITHMS = {"direct-sum": DIRECT_SUM, "barnes-hut": BARNES_HUT}

DUMP_CELLS, DUMP_MORTON30, DUMP_SORTED_SRC = 1, 2, 3
DUMP_COUNTS, DUMP_RADII, DUMP_NODE_POS, DUMP_NODE_MASS = 4, 5, 6, 7
DUMP_NODE_TOTAL, DUMP_ACCEL, DUMP_COUNTERS, DUMP_BOX = 8, 9, 10, 11
DEBUG_ACCEL = 1
TUNE_DS_RCP, TUNE_BH_WAVES_PER_SIMD, TUNE_BH_TRAVERSAL = 1, 2, 3
TUNE_DS_SYMMETRIC, TUNE_DS_SYM_COLS, TUNE_BH_SEGMENTS, TUNE_BH_SPLIT = 4, 5, 6, 7
TUNE_BH_ADAPTIVE_BOX = 8
TUNE_BH_LEAF_BITS = 9
TUNE_BH_XCD_ORDER = 10
TUNE_BH_ORDER = 11

STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE",
          4: "NO_DEVICE", 5: "IO", 6: "CAPACITY"}


class GsimError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("gravitational_const", C.c_float),
        ("simulation_time", C.c_float),
        ("simulation_speed", C.c_float),
        ("softening_len", C.c_float),
        ("accuracy_parameter", C.c_float),
        ("algorithm", C.c_int32),
        ("group_width", C.c_uint32),
        ("device", C.c_int32),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("force_ms", C.c_double),
        ("build_ms", C.c_double),
        ("integrate_ms", C.c_double),
        ("step_ms", C.c_double),
        ("steps", C.c_uint64),
        ("force_launches", C.c_uint64),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("systems", C.c_uint32), ("max_systems", C.c_uint32),
        ("capacity_particles", C.c_uint64), ("used_particles", C.c_uint64),
        ("largest_free_range", C.c_uint64), ("node_capacity", C.c_uint64),
        ("nodes_reserved", C.c_uint64), ("steps", C.c_uint64), ("waves_per_simd", C.c_uint32),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("waves_per_simd", C.c_uint32), ("sort_tile", C.c_uint32),
                ("run_queue", C.c_uint32), ("prio_step", C.c_uint32), ("prio_carry", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("steps", C.c_uint64),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64), ("failed", C.c_uint64), ("queued", C.c_uint64),
        ("running", C.c_uint64), ("batch_steps", C.c_uint64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
        ("span_s", C.c_double), ("ds_interactions", C.c_uint64),
        ("bh_interactions", C.c_uint64), ("capacity_particles", C.c_uint64),
        ("bh_visits", C.c_uint64), ("arenas", C.c_uint32),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libgsim.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GsimError(-1, f"{path} not found: build it with python gravity-sim_amd/build.py")
    lib = C.CDLL(path)
    fp = C.POINTER(C.c_float)
    vp = C.c_void_p
    st = C.c_int
    sig = {
        "gsim_default_params": (None, [C.POINTER(Params)]),
        "gsim_last_error": (C.c_char_p, []),
        "gsim_version": (C.c_char_p, []),
        "gsim_required_alignment": (C.c_size_t, [C.c_int32]),
        "gsim_create": (st, [C.POINTER(Params), fp, fp, fp, C.c_size_t, C.POINTER(vp)]),
        "gsim_step": (st, [vp, C.c_uint32]),
        "gsim_synchronize": (st, [vp]),
        "gsim_read": (st, [vp, fp, fp, fp]),
        "gsim_destroy": (st, [vp]),
        "gsim_particle_count": (C.c_size_t, [vp]),
        "gsim_aligned_particle_count": (C.c_size_t, [vp]),
        "gsim_steps_done": (C.c_uint64, [vp]),
        "gsim_debug_dump": (st, [vp, C.c_int32, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "gsim_set_debug": (st, [vp, C.c_uint32]),
        "gsim_set_timing": (st, [vp, C.c_int32]),
        "gsim_get_timing": (st, [vp, C.POINTER(Timing), C.c_int32]),
        "gsim_set_ds_shape": (st, [vp, C.c_uint32, C.c_uint32]),
        "gsim_set_tuning": (st, [vp, C.c_int32, C.c_uint32]),
        "gsim_batch_create": (st, [C.c_int32, C.c_size_t, C.c_uint32, C.POINTER(vp)]),
        "gsim_batch_add": (st, [vp, C.POINTER(Params), fp, fp, fp, C.c_size_t,
                                C.POINTER(C.c_uint32)]),
        "gsim_batch_remove": (st, [vp, C.c_uint32]),
        "gsim_batch_step": (st, [vp, C.c_uint32]),
        "gsim_batch_synchronize": (st, [vp]),
        "gsim_batch_read": (st, [vp, C.c_uint32, fp, fp, fp]),
        "gsim_batch_query": (st, [vp, C.POINTER(BatchInfo)]),
        "gsim_batch_bytes_per_particle": (C.c_size_t, []),
        "gsim_batch_set_waves": (st, [vp, C.c_uint32]),
        "gsim_batch_autotune": (st, [vp, C.c_uint32, C.POINTER(C.c_uint32)]),
        "gsim_batch_get_tuning": (st, [vp, C.POINTER(BatchTuning)]),
        "gsim_batch_set_tuning": (st, [vp, C.POINTER(BatchTuning)]),
        "gsim_batch_destroy": (st, [vp]),
        "gsim_batch_set_steps": (st, [vp, C.c_uint32, C.c_uint64]),
        "gsim_batch_steps_done": (st, [vp, C.c_uint32, C.POINTER(C.c_uint64)]),
        "gsim_batch_accepts": (st, [vp, C.POINTER(C.c_uint64)]),
        "gsim_batch_visits": (st, [vp, C.POINTER(C.c_uint64)]),
        "gsim_server_create": (st, [C.c_int32, C.c_size_t, C.c_uint32, C.c_size_t,
                                    C.POINTER(vp)]),
        "gsim_server_create_arenas": (st, [C.c_int32, C.c_size_t, C.c_uint32, C.c_size_t,
                                           C.c_size_t, C.POINTER(vp)]),
        "gsim_server_submit": (st, [vp, C.POINTER(Params), fp, fp, fp, C.c_size_t,
                                    C.c_uint32, C.POINTER(C.c_uint64)]),
        "gsim_server_wait": (st, [vp, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
        "gsim_server_result": (st, [vp, C.c_uint64, fp, fp, fp, C.POINTER(RequestInfo)]),
        "gsim_server_release": (st, [vp, C.c_uint64]),
        "gsim_server_drain": (st, [vp, C.c_double, C.POINTER(C.c_int32)]),
        "gsim_server_query": (st, [vp, C.POINTER(ServerInfo)]),
        "gsim_server_destroy": (st, [vp]),
        "gsim_generate": (st, [C.c_int32, C.c_size_t, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_uint64, fp, fp, fp]),
        "gsim_count_particles": (st, [C.c_char_p, C.POINTER(C.c_size_t)]),
        "gsim_load_particles": (st, [C.c_char_p, fp, fp, fp, C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
        "gsim_save_particles": (st, [C.c_char_p, fp, fp, fp, C.c_size_t]),
        "gsim_ds_step_noalloc": (st, [C.POINTER(Params), vp, vp, vp, vp, vp, C.c_size_t,
                                      vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = _lib.gsim_last_error().decode(errors="replace") if _lib else ""
        raise GsimError(status, msg)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def default_params() -> Params:
    p = Params()
    load_library().gsim_default_params(C.byref(p))
    return p


def ds_step_noalloc(params: Params, pos_xy: int, vel_xy: int, mass: int, pos_out_xy: int,
                    vel_out_xy: int, n: int, stream: int | None = None) -> None:
    """One direct-sum step on caller-owned device arrays (addresses as ints),
    with no device allocation: gsim_ds_step_noalloc.  Asynchronous on
    `stream` (a hipStream_t address; None = the null stream)."""
    _check(load_library().gsim_ds_step_noalloc(C.byref(params), pos_xy, vel_xy, mass,
                                               pos_out_xy, vel_out_xy, n, stream))


def required_alignment(algorithm: int) -> int:
    return int(load_library().gsim_required_alignment(algorithm))


@dataclass
class Particles:
    """Host particle set: pos (n,2), vel (n,2), mass (n,), all float32."""
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray

    @property
    def n(self) -> int:
        return int(self.mass.shape[0])

    def copy(self) -> "Particles":
        return Particles(self.pos.copy(), self.vel.copy(), self.mass.copy())

    def padded(self, alignment: int = 64) -> "Particles":
        na = (self.n + alignment - 1) // alignment * alignment
        p = np.zeros((na, 2), np.float32)
        v = np.zeros((na, 2), np.float32)
        m = np.zeros(na, np.float32)
        p[: self.n] = self.pos
        v[: self.n] = self.vel
        m[: self.n] = self.mass
        return Particles(p, v, m)


def generate(kind, n: int, generate_size: float, min_mass: float, max_mass: float,
             gravitational_const: float = 1.0, seed: int = 0) -> Particles:
    """One of the reference presets (or a synthetic disk), seeded."""
    lib = load_library()
    t = GENERATE_TYPES[kind] if isinstance(kind, str) else int(kind)
    pos = np.empty((n, 2), np.float32)
    vel = np.empty((n, 2), np.float32)
    mass = np.empty(n, np.float32)
    _check(lib.gsim_generate(t, n, generate_size, min_mass, max_mass, gravitational_const,
                             seed, _fptr(pos), _fptr(vel), _fptr(mass)))
    return Particles(pos, vel, mass)


def load_particles(path: str) -> Particles:
    lib = load_library()
    n = C.c_size_t(0)
    _check(lib.gsim_count_particles(path.encode(), C.byref(n)))
    cnt = n.value
    pos = np.empty((cnt, 2), np.float32)
    vel = np.empty((cnt, 2), np.float32)
    mass = np.empty(cnt, np.float32)
    _check(lib.gsim_load_particles(path.encode(), _fptr(pos), _fptr(vel), _fptr(mass), cnt,
                                   C.byref(n)))
    return Particles(pos, vel, mass)


def save_particles(path: str, p: Particles) -> None:
    lib = load_library()
    pos = np.ascontiguousarray(p.pos, np.float32)
    vel = np.ascontiguousarray(p.vel, np.float32)
    mass = np.ascontiguousarray(p.mass, np.float32)
    _check(lib.gsim_save_particles(path.encode(), _fptr(pos), _fptr(vel), _fptr(mass), p.n))


class Simulation:
    """One device-resident simulation (a gsim_sim handle)."""

    def __init__(self, particles: Particles, algorithm: int | str = DIRECT_SUM, *,
                 gravitational_const: float = 1.0, simulation_time: float = 1e-3,
                 simulation_speed: float = 1.0, softening_len: float = 0.2,
                 accuracy_parameter: float = 1.0, group_width: int = 64, device: int = 0):
        lib = load_library()
        p = Params()
        lib.gsim_default_params(C.byref(p))
        p.gravitational_const = gravitational_const
        p.simulation_time = simulation_time
        p.simulation_speed = simulation_speed
        p.softening_len = softening_len
        p.accuracy_parameter = accuracy_parameter
        p.algorithm = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        p.group_width = group_width
        p.device = device
        self.params = p
        pos = np.ascontiguousarray(particles.pos, np.float32)
        vel = np.ascontiguousarray(particles.vel, np.float32)
        mass = np.ascontiguousarray(particles.mass, np.float32)
        h = C.c_void_p()
        _check(lib.gsim_create(C.byref(p), _fptr(pos), _fptr(vel), _fptr(mass),
                               particles.n, C.byref(h)))
        self._h = h
        self.n = particles.n
        self.na = int(lib.gsim_aligned_particle_count(h))

    # RunSimulations(n): asynchronous, stream ordered
    def step(self, count: int = 1) -> None:
        _check(_lib.gsim_step(self._h, count))

    run_simulations = step

    def synchronize(self) -> None:
        _check(_lib.gsim_synchronize(self._h))

    def read(self) -> Particles:
        pos = np.empty((self.n, 2), np.float32)
        vel = np.empty((self.n, 2), np.float32)
        mass = np.empty(self.n, np.float32)
        _check(_lib.gsim_read(self._h, _fptr(pos), _fptr(vel), _fptr(mass)))
        return Particles(pos, vel, mass)

    def set_debug(self, flags: int) -> None:
        _check(_lib.gsim_set_debug(self._h, flags))

    def set_timing(self, enable: bool) -> None:
        _check(_lib.gsim_set_timing(self._h, 1 if enable else 0))

    def timing(self, reset: bool = True) -> dict:
        t = Timing()
        _check(_lib.gsim_get_timing(self._h, C.byref(t), 1 if reset else 0))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def set_ds_shape(self, targets_per_lane: int, slices: int) -> None:
        _check(_lib.gsim_set_ds_shape(self._h, targets_per_lane, slices))

    def set_tuning(self, knob: int, value: int) -> None:
        _check(_lib.gsim_set_tuning(self._h, knob, value))

    def dump(self, what: int) -> np.ndarray:
        written = C.c_size_t(0)
        status = _lib.gsim_debug_dump(self._h, what, None, 0, C.byref(written))
        if status not in (0, 6):
            _check(status)
        nbytes = written.value
        dtype = {DUMP_RADII: np.float32, DUMP_NODE_POS: np.float32,
                 DUMP_NODE_MASS: np.float32, DUMP_ACCEL: np.float32,
                 DUMP_COUNTERS: np.uint64, DUMP_BOX: np.float32}.get(what, np.uint32)
        buf = np.zeros(max(nbytes, 1) // np.dtype(dtype).itemsize, dtype)
        _check(_lib.gsim_debug_dump(self._h, what, buf.ctypes.data_as(C.c_void_p),
                                    buf.nbytes, C.byref(written)))
        buf = buf[: written.value // np.dtype(dtype).itemsize]
        if what in (DUMP_NODE_POS, DUMP_ACCEL):
            buf = buf.reshape(-1, 2)
        return buf

    def tree(self) -> dict:
        """Node arrays of the last Barnes-Hut build in the reference's layout."""
        return {
            "counts": self.dump(DUMP_COUNTS),
            "radii": self.dump(DUMP_RADII),
            "node_pos": self.dump(DUMP_NODE_POS),
            "node_mass": self.dump(DUMP_NODE_MASS),
            "sorted_src": self.dump(DUMP_SORTED_SRC),
            "total": int(self.dump(DUMP_NODE_TOTAL)[0]),
        }

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.gsim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_params(algorithm=DIRECT_SUM, *, gravitational_const: float = 1.0,
                simulation_time: float = 1e-3, simulation_speed: float = 1.0,
                softening_len: float = 0.2, accuracy_parameter: float = 1.0,
                group_width: int = 64, device: int = 0) -> Params:
    p = default_params()
    p.gravitational_const = gravitational_const
    p.simulation_time = simulation_time
    p.simulation_speed = simulation_speed
    p.softening_len = softening_len
    p.accuracy_parameter = accuracy_parameter
    p.algorithm = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    p.group_width = group_width
    p.device = device
    return p


class Batch:
    """Many independent systems in one device arena, stepped together
    (gsim_batch_*).  add() returns a system id; step() advances every system."""

    def __init__(self, capacity_particles: int, max_systems: int = 256, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.gsim_batch_create(device, capacity_particles, max_systems, C.byref(h)))
        self._h = h
        self.n = {}

    @staticmethod
    def bytes_per_particle() -> int:
        return int(load_library().gsim_batch_bytes_per_particle())

    def add(self, particles: Particles, algorithm=DIRECT_SUM, **kw) -> int:
        p = make_params(algorithm, **kw)
        pos = np.ascontiguousarray(particles.pos, np.float32)
        vel = np.ascontiguousarray(particles.vel, np.float32)
        mass = np.ascontiguousarray(particles.mass, np.float32)
        sid = C.c_uint32(0)
        _check(_lib.gsim_batch_add(self._h, C.byref(p), _fptr(pos), _fptr(vel), _fptr(mass),
                                   particles.n, C.byref(sid)))
        self.n[sid.value] = particles.n
        return sid.value

    def remove(self, sid: int) -> None:
        _check(_lib.gsim_batch_remove(self._h, sid))
        self.n.pop(sid, None)

    def step(self, count: int = 1) -> None:
        _check(_lib.gsim_batch_step(self._h, count))

    def synchronize(self) -> None:
        _check(_lib.gsim_batch_synchronize(self._h))

    def read(self, sid: int) -> Particles:
        n = self.n[sid]
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        mass = np.empty(n, np.float32)
        _check(_lib.gsim_batch_read(self._h, sid, _fptr(pos), _fptr(vel), _fptr(mass)))
        return Particles(pos, vel, mass)

    def info(self) -> dict:
        i = BatchInfo()
        _check(_lib.gsim_batch_query(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}

    def set_waves(self, waves_per_simd: int) -> None:
        _check(_lib.gsim_batch_set_waves(self._h, waves_per_simd))

    def set_steps(self, sid: int, more_steps: int | None) -> None:
        """Limits a system to more_steps further steps (None: unlimited)."""
        v = (1 << 64) - 1 if more_steps is None else int(more_steps)
        _check(_lib.gsim_batch_set_steps(self._h, sid, v))

    def steps_done(self, sid: int) -> int:
        v = C.c_uint64(0)
        _check(_lib.gsim_batch_steps_done(self._h, sid, C.byref(v)))
        return v.value

    def accepts(self) -> int:
        v = C.c_uint64(0)
        _check(_lib.gsim_batch_accepts(self._h, C.byref(v)))
        return v.value

    def visits(self) -> int:
        v = C.c_uint64(0)
        _check(_lib.gsim_batch_visits(self._h, C.byref(v)))
        return v.value

    def autotune(self, steps_per_trial: int = 5) -> int:
        """Tunes every launch knob (see tuning()); returns the chosen waves
        per SIMD."""
        w = C.c_uint32(0)
        _check(_lib.gsim_batch_autotune(self._h, steps_per_trial, C.byref(w)))
        return w.value

    def tuning(self) -> dict:
        t = BatchTuning()
        _check(_lib.gsim_batch_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_}

    def set_tuning(self, **kw) -> None:
        """Sets any of waves_per_simd, sort_tile, run_queue, prio_step,
        prio_carry; the others keep their values."""
        cur = self.tuning()
        cur.update(kw)
        t = BatchTuning(**cur)
        _check(_lib.gsim_batch_set_tuning(self._h, C.byref(t)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.gsim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching server (gsim_server_*): submit() queues a system for
    a number of steps and returns a request id; a worker thread in the library
    admits requests into one batch as HBM allows and retires each after exactly
    its steps.  result() returns the final particles and the request's
    latency."""

    def __init__(self, hbm_budget_bytes: int, max_systems: int = 1024,
                 max_active_particles: int = 0, device: int = 0, arena_particles: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.gsim_server_create_arenas(device, hbm_budget_bytes, max_systems,
                                             max_active_particles, arena_particles,
                                             C.byref(h)))
        self._h = h
        self.n = {}

    def submit(self, particles: Particles, steps: int, algorithm=DIRECT_SUM, **kw) -> int:
        p = make_params(algorithm, **kw)
        pos = np.ascontiguousarray(particles.pos, np.float32)
        vel = np.ascontiguousarray(particles.vel, np.float32)
        mass = np.ascontiguousarray(particles.mass, np.float32)
        rid = C.c_uint64(0)
        _check(_lib.gsim_server_submit(self._h, C.byref(p), _fptr(pos), _fptr(vel), _fptr(mass),
                                       particles.n, steps, C.byref(rid)))
        self.n[rid.value] = particles.n
        return rid.value

    def wait(self, rid: int, timeout_s: float = 60.0) -> bool:
        done = C.c_int32(0)
        _check(_lib.gsim_server_wait(self._h, rid, timeout_s, C.byref(done)))
        return bool(done.value)

    def result(self, rid: int, release: bool = True) -> tuple[Particles, dict]:
        if rid not in self.n:
            raise GsimError(1, f"unknown request {rid}")
        n = self.n[rid]
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        mass = np.empty(n, np.float32)
        info = RequestInfo()
        _check(_lib.gsim_server_result(self._h, rid, _fptr(pos), _fptr(vel), _fptr(mass),
                                       C.byref(info)))
        if release:
            _check(_lib.gsim_server_release(self._h, rid))
            self.n.pop(rid, None)
        return Particles(pos, vel, mass), {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def drain(self, timeout_s: float = 600.0) -> bool:
        idle = C.c_int32(0)
        _check(_lib.gsim_server_drain(self._h, timeout_s, C.byref(idle)))
        return bool(idle.value)

    def stats(self) -> dict:
        i = ServerInfo()
        _check(_lib.gsim_server_query(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in ServerInfo._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.gsim_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ParticleSystem:
    """Particle set plus parameters, the reference's ParticleSystem surface
    (src/Particles/ParticleSystem.hpp:58-123): built either from a particle file
    or from a generation preset, padded to the algorithm's alignment."""

    def __init__(self, particles: Particles, gravitational_const: float = 1.0,
                 simulation_time: float = 1e-3, simulation_speed: float = 1.0,
                 softening_len: float = 0.2, accuracy_parameter: float = 1.0,
                 simulation_algorithm: int | str = DIRECT_SUM):
        self.particles = particles
        self.gravitational_const = gravitational_const
        self.simulation_time = simulation_time
        self.simulation_speed = simulation_speed
        self.softening_len = softening_len
        self.accuracy_parameter = accuracy_parameter
        self.simulation_algorithm = (ALGORITHMS[simulation_algorithm]
                                     if isinstance(simulation_algorithm, str)
                                     else int(simulation_algorithm))
        align = required_alignment(self.simulation_algorithm)
        self.particle_count = particles.n
        self.aligned_particle_count = (particles.n + align - 1) // align * align

    @classmethod
    def from_file(cls, path: str, **kw) -> "ParticleSystem":
        return cls(load_particles(path), **kw)

    @classmethod
    def generated(cls, count: int, generate_type, generate_size: float, min_mass: float,
                  max_mass: float, seed: int = 0, **kw) -> "ParticleSystem":
        t = GENERATE_TYPES[generate_type] if isinstance(generate_type, str) else generate_type
        if t == GENERATE_SYMMETRICAL_GALAXY_COLLISION:
            count &= ~1  # ParticleSystem.cpp:522-525
        g = kw.get("gravitational_const", 1.0)
        return cls(generate(t, count, generate_size, min_mass, max_mass, g, seed), **kw)


class _SimulationBase:
    algorithm = DIRECT_SUM

    def __init__(self, particle_system: ParticleSystem, group_width: int = 64, device: int = 0):
        ps = particle_system
        self.particle_system = ps
        self.sim = Simulation(ps.particles, self.algorithm,
                              gravitational_const=ps.gravitational_const,
                              simulation_time=ps.simulation_time,
                              simulation_speed=ps.simulation_speed,
                              softening_len=ps.softening_len,
                              accuracy_parameter=ps.accuracy_parameter,
                              group_width=group_width, device=device)

    @staticmethod
    def get_required_particle_alignment() -> int:
        return 64

    def run_simulations(self, simulation_count: int) -> None:
        self.sim.step(simulation_count)

    def get_particles(self) -> Particles:
        return self.sim.read()

    def save_particles(self, path: str) -> None:
        save_particles(path, self.sim.read())

    def close(self) -> None:
        self.sim.close()


class DirectSimulation(_SimulationBase):
    algorithm = DIRECT_SUM


class BarnesHutSimulation(_SimulationBase):
    algorithm = BARNES_HUT
