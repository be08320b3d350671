"""CPU tests of libgsim's C ABI: it loads, exports every declared symbol, and
its host-side surface (presets, particle text format, argument checks) works
without a GPU.  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import gsim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^[A-Za-z_][\w \*]*?\b(gsim_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = gsim.load_library()
    decl = _declared_functions()
    assert len(decl) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", gsim.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gsim_\w+)$", out, flags=re.M))
    missing = [n for n in decl if n not in exported]
    assert not missing, missing
    for n in decl:
        assert getattr(lib, n) is not None


def test_defaults_and_alignment_match_reference():
    p = gsim.default_params()
    # Main.cpp:59-63
    assert (p.gravitational_const, p.simulation_speed, p.accuracy_parameter) == (1.0, 1.0, 1.0)
    assert p.simulation_time == np.float32(1e-3) and p.softening_len == np.float32(0.2)
    assert gsim.required_alignment(gsim.DIRECT_SUM) == 64
    assert gsim.required_alignment(gsim.BARNES_HUT) == 64
    assert gsim.required_alignment(7) == 0


@pytest.mark.parametrize("kind", ["random", "galaxy", "galaxy-collision",
                                  "symmetrical-galaxy-collision", "exponential-disk",
                                  "two-exponential-disks"])
def test_presets_are_seeded_and_deterministic(kind):
    a = gsim.generate(kind, 1000, 100.0, 1.0, 10.0, seed=5)
    b = gsim.generate(kind, 1000, 100.0, 1.0, 10.0, seed=5)
    c = gsim.generate(kind, 1000, 100.0, 1.0, 10.0, seed=6)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    assert not np.array_equal(a.pos, c.pos)
    assert np.all((a.mass >= 1) & (a.mass <= 10))
    assert np.isfinite(a.pos).all() and np.isfinite(a.vel).all()


def test_random_preset_geometry():
    # ParticleSystem.cpp:15-35: r = sqrt(1-u)*size (uniform in area), v = 0
    p = gsim.generate("random", 20000, 100.0, 1.0, 10.0, seed=1)
    r = np.linalg.norm(p.pos, axis=1)
    assert r.max() <= 100.0 + 1e-3
    assert np.all(p.vel == 0)
    # uniform in area: P(r < R/2) = 1/4
    assert abs((r < 50).mean() - 0.25) < 0.01


def test_galaxy_preset_velocity_formula():
    # ParticleSystem.cpp:36-62: v = (-sin t, cos t) * r * sqrt(G N avg / size^3)
    n, size, G = 5000, 100.0, 1.0
    p = gsim.generate("galaxy", n, size, 1.0, 10.0, G, seed=2)
    r = np.linalg.norm(p.pos.astype(np.float64), axis=1)
    vel = np.sqrt(G * n * 5.5 / size ** 3)
    assert np.allclose(np.linalg.norm(p.vel, axis=1), r * vel, rtol=1e-5, atol=1e-5)
    assert np.allclose((p.pos * p.vel).sum(1), 0, atol=1e-2)  # tangential


def test_collision_presets_layout():
    n, size = 4000, 90.0
    p = gsim.generate("galaxy-collision", n, size, 1.0, 10.0, seed=3)
    gs = size / 3
    left, right = p.pos[: n // 2], p.pos[n // 2:]
    assert np.all(np.linalg.norm(left - [-2 * gs, 0], axis=1) <= gs + 1e-3)
    assert np.all(np.linalg.norm(right - [2 * gs, 0], axis=1) <= gs + 1e-3)
    s = gsim.generate("symmetrical-galaxy-collision", n, size, 1.0, 10.0, seed=3)
    assert np.array_equal(s.pos[1::2, 0], -s.pos[0::2, 0])
    assert np.array_equal(s.pos[1::2, 1], s.pos[0::2, 1])
    assert np.array_equal(s.mass[1::2], s.mass[0::2])
    with pytest.raises(gsim.GsimError):
        gsim.generate("symmetrical-galaxy-collision", 3, size, 1.0, 10.0)


def test_two_disk_workload_geometry():
    p = gsim.generate("two-exponential-disks", 20000, 33.0, 1.0, 10.0, seed=4)
    left, right = p.pos[:10000], p.pos[10000:]
    assert np.all(np.linalg.norm(left - [-40, 0], axis=1) <= 33.0 + 1e-3)
    assert np.all(np.linalg.norm(right - [40, 0], axis=1) <= 33.0 + 1e-3)
    # bulk velocities +2 and -2 on top of purely circular internal rotation:
    # after removing the bulk, velocity is perpendicular to the radius vector
    for part, c, bulk in ((left, -40.0, 2.0), (right, 40.0, -2.0)):
        sl = slice(0, 10000) if bulk > 0 else slice(10000, 20000)
        rel_v = p.vel[sl].astype(np.float64) - [bulk, 0.0]
        rel_p = part.astype(np.float64) - [c, 0.0]
        cosang = (rel_v * rel_p).sum(1) / (np.linalg.norm(rel_v, axis=1) *
                                           np.linalg.norm(rel_p, axis=1) + 1e-12)
        assert np.abs(cosang).max() < 1e-3
        assert (rel_p[:, 0] * rel_v[:, 1] - rel_p[:, 1] * rel_v[:, 0]).min() >= 0  # CCW
    # exponential disk with h = 7 truncated at 33: radius ~ Gamma(2, 7) | r <= 33,
    # CDF F(x) = 1 - (1 + x/h) exp(-x/h); compare the sample median with it
    xs = np.linspace(0, 33, 330001)
    F = 1 - (1 + xs / 7.0) * np.exp(-xs / 7.0)
    med_ref = xs[np.searchsorted(F / F[-1], 0.5)]
    med = np.median(np.linalg.norm(left - [-40, 0], axis=1))
    assert abs(med - med_ref) < 0.3


def test_particle_file_round_trip(tmp_path):
    p = gsim.generate("galaxy", 257, 50.0, 1.0, 10.0, seed=9)
    path = str(tmp_path / "p.txt")
    gsim.save_particles(path, p)
    lines = open(path).read().splitlines()
    assert len(lines) == 257
    # "%.7f %.7f %.7f %.7f %.7f" (ParticleSystem.cpp:743)
    assert re.fullmatch(r"(-?\d+\.\d{7} ){4}-?\d+\.\d{7}", lines[0])
    q = gsim.load_particles(path)
    assert q.n == 257
    assert np.allclose(q.pos, p.pos, atol=1e-6) and np.allclose(q.mass, p.mass, atol=1e-6)


def test_particle_file_parsing_stops_at_first_bad_record(tmp_path):
    path = tmp_path / "p.txt"
    path.write_text("1 2 3 4 5\n6 7 8 9 10\n1 2 3 oops 5\n1 1 1 1 1\n")
    q = gsim.load_particles(str(path))
    assert q.n == 2 and q.mass.tolist() == [5.0, 10.0]
    empty = tmp_path / "e.txt"
    empty.write_text("nothing here\n")
    with pytest.raises(gsim.GsimError) as e:
        gsim.load_particles(str(empty))
    assert "at least one valid particle" in str(e.value)
    with pytest.raises(gsim.GsimError):
        gsim.load_particles(str(tmp_path / "missing.txt"))


def test_create_rejects_bad_arguments_before_touching_a_device():
    p = gsim.generate("random", 10, 10.0, 1.0, 2.0)
    for kw in ({"algorithm": 5}, {"group_width": 48}, {"accuracy_parameter": float("nan")},
               {"softening_len": float("inf")}):
        with pytest.raises(gsim.GsimError) as e:
            gsim.Simulation(p, **kw)
        assert e.value.status == 1
    empty = gsim.Particles(np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32),
                           np.zeros(0, np.float32))
    with pytest.raises(gsim.GsimError) as e:
        gsim.Simulation(empty)
    assert "at least one particle" in str(e.value)


def test_no_cpu_fallback_without_device():
    """Without a gfx950 device the library refuses to run instead of falling
    back to a CPU path."""
    p = gsim.generate("random", 10, 10.0, 1.0, 2.0)
    try:
        sim = gsim.Simulation(p)
    except gsim.GsimError as e:
        assert e.status == 4  # NO_DEVICE
        return
    sim.close()
    pytest.skip("a GPU is visible; covered by the gpu tests")


def test_python_constants_mirror_the_header():
    """Every GSIM_TUNE_*, GSIM_DUMP_* and GSIM_DEBUG_* value in include/gsim.h
    has a Python constant of the same value (the ctypes layer passes these
    integers straight to the C ABI), and the Python side defines no value the
    header lacks."""
    text = open(os.path.join(ROOT, "include", "gsim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    header = {}
    for name, val in re.findall(r"\bGSIM_((?:TUNE|DUMP)_\w+)\s*=\s*(\d+)", text):
        header[name] = int(val)
    for name, val in re.findall(r"#define\s+GSIM_(DEBUG_\w+)\s+(\d+)u?", text):
        header[name] = int(val)
    assert len([k for k in header if k.startswith("TUNE_")]) >= 11
    assert len([k for k in header if k.startswith("DUMP_")]) >= 11
    for name, val in header.items():
        assert getattr(gsim, name) == val, name
    py = {k for k in dir(gsim) if re.fullmatch(r"(TUNE|DUMP|DEBUG)_[A-Z0-9_]+", k)}
    assert py == set(header), sorted(py ^ set(header))
