"""The gravity-sim command line (the reference's headless Main.cpp surface).

This file is named to run first: it starts the CLI as a child process, which
must happen before this pytest process makes any HIP call.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import gsim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gravity-sim_amd", "bin", "gravity-sim")


def run(*args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_help_prints_and_exits():
    r = run("--help")
    assert r.returncode == 0
    assert "--simulation-algorithm" in r.stdout and "--particles-out" in r.stdout
    # every parameter and option of the reference's help (Main.cpp:20-48),
    # each generate type and algorithm with its own description line
    for flag in ("--log-file:", "--particles-in:", "--particles-out:", "--particle-count:",
                 "--generate-type:", "--generate-size:", "--min-mass:", "--max-mass:",
                 "--gravitational-const:", "--simulation-time:", "--simulation-speed:",
                 "--softening-len:", "--accuracy-parameter:", "--simulation-algorithm:",
                 "--simulation-count:", "--help:", "--log-detailed:", "--no-graphics:",
                 "--benchmark:"):
        assert "\t" + flag in r.stdout, flag
    for choice in ("random", "galaxy", "galaxy-collision", "symmetrical-galaxy-collision",
                   "direct-sum", "barnes-hut"):
        assert re.search(r"^\t\t" + re.escape(choice) + r": \S", r.stdout, re.M), choice


def test_readme_example_arguments_parse():
    """The README's command line passes every valued argument in the
    reference's --key=value form (Main.cpp:168-238), which is the only form
    the CLI parses."""
    readme = open(os.path.join(ROOT, "README.md")).read()
    m = re.search(r"gravity-sim_amd/bin/gravity-sim (.*?)\n\n", readme, re.S)
    assert m, "README example not found"
    args = m.group(1).replace("\\\n", " ").split()
    assert all(a.startswith("--") for a in args)
    valued = [a for a in args if a not in ("--no-graphics", "--benchmark")]
    assert all("=" in a for a in valued), valued


@pytest.mark.parametrize("args,msg", [
    ([], "If a particle input file is not provided, valid generation args must be provided!"),
    (["--particle-count=10", "--generate-type=random", "--generate-size=5", "--min-mass=1",
      "--max-mass=2"], "A valid simulation algorithm must be given!"),
    (["--particle-count=10", "--generate-type=random", "--generate-size=5", "--min-mass=1",
      "--max-mass=2", "--simulation-algorithm=direct-sum", "--no-graphics"],
     "If --no-graphics was specified, a valid simulation count must be given!"),
    (["--particle-count=10", "--generate-type=nope", "--generate-size=5", "--min-mass=1",
      "--max-mass=2", "--simulation-algorithm=direct-sum"],
     "If a particle input file is not provided"),
])
def test_invalid_arguments_are_fatal(args, msg):
    r = run(*args)
    assert r.returncode != 0  # FATAL aborts, Logger.cpp:46-47
    assert "[FATAL ERROR] " + msg in r.stdout


def test_log_file_and_generation_warning(tmp_path):
    pfile = tmp_path / "p.txt"
    pfile.write_text("1 2 0 0 1\n")
    log = tmp_path / "log.txt"
    r = run(f"--particles-in={pfile}", "--particle-count=5", f"--log-file={log}",
            "--simulation-algorithm=barnes-hut")
    assert "[WARNING]     A particle input file was provided" in r.stdout
    assert "[WARNING]     A particle input file was provided" in log.read_text()


def _final_state(algo, pin, steps, extra=(), tunings=()):
    p = gsim.load_particles(str(pin))
    with gsim.Simulation(p, algo, **dict(extra)) as s:
        for knob, value in tunings:
            s.set_tuning(knob, value)
        s.step(steps)
        return s.read()


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["direct-sum", "barnes-hut"])
def test_cli_run_matches_library(tmp_path, algo):
    """The CLI (file in, 250 steps in batches of 100, file out) produces the
    same final state as the library driven directly from the same file."""
    pin = tmp_path / "in.txt"
    pout = tmp_path / "out.txt"
    gsim.save_particles(str(pin), gsim.generate("galaxy-collision", 3000, 150.0, 1.0, 10.0,
                                                seed=7))
    r = run(f"--particles-in={pin}", f"--particles-out={pout}", f"--simulation-algorithm={algo}",
            "--simulation-count=250", "--no-graphics", "--benchmark", "--simulation-speed=2")
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"\[INFO\]        Total simulation runtime: [\d.]+(ms|s)", r.stdout)
    assert re.search(r"\[INFO\]        Average runtime/simulation: [\d.]+(us|ms|s)", r.stdout)
    got = gsim.load_particles(str(pout))
    # the library run is started only now, after the child process has exited
    ref = _final_state(algo, pin, 250, {"simulation_speed": 2.0})
    assert got.n == 3000
    # the file holds %.7f, so compare at that resolution
    assert np.allclose(got.pos, ref.pos, atol=1e-6, rtol=0)
    assert np.allclose(got.vel, ref.vel, atol=1e-6, rtol=0)
    assert np.allclose(got.mass, ref.mass, atol=1e-6, rtol=0)


@pytest.mark.gpu
def test_cli_tree_options_match_library(tmp_path):
    """--leaf-bits=15 --adaptive-box (beyond the reference): the CLI's final
    state equals the library's with the same two options; a leaf-bit count
    other than 10 or 15 is refused."""
    pin = tmp_path / "in.txt"
    pout = tmp_path / "out.txt"
    gsim.save_particles(str(pin), gsim.generate("galaxy-collision", 3000, 700.0, 1.0, 10.0,
                                                seed=8))
    r = run(f"--particles-in={pin}", f"--particles-out={pout}", "--simulation-algorithm=barnes-hut",
            "--simulation-count=120", "--no-graphics", "--leaf-bits=15", "--adaptive-box")
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    bad = run(f"--particles-in={pin}", "--simulation-algorithm=barnes-hut",
              "--simulation-count=1", "--no-graphics", "--leaf-bits=12")
    assert bad.returncode != 0 and "leaf bits must be 10 or 15" in bad.stdout + bad.stderr
    got = gsim.load_particles(str(pout))
    ref = _final_state("barnes-hut", pin, 120, tunings=((gsim.TUNE_BH_LEAF_BITS, 15),
                                                       (gsim.TUNE_BH_ADAPTIVE_BOX, 1)))
    assert np.array_equal(ref.mass, gsim.load_particles(str(pin)).mass)  # nothing killed
    assert np.allclose(got.pos, ref.pos, atol=1e-6, rtol=0)
    assert np.allclose(got.vel, ref.vel, atol=1e-6, rtol=0)
