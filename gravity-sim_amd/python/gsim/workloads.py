"""The benchmark workloads as seeded generators, shared by bench.py and the
GPU tests so that the systems a test checks are the systems the bench times.

    direct_sum_disk()      the 50,000-particle exponential disk (direct sum)
    two_disks()            the 500,000-particle pair of approaching disks
                           (Barnes-Hut, theta = 1)
    batched_systems()      256 independent systems of uneven size
    served_requests()      a Poisson request stream of mixed systems

Every generator takes an explicit seed; the presets are the reference's
(ParticleSystem.cpp:15-149) with a SplitMix64 stream in place of its
time-seeded engine.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Particles, generate

DS_N = 50_000
BH_N = 500_000


def direct_sum_disk(seed: int = 2) -> Particles:
    """One exponential disk: radius 100, scale length 20, circular velocities
    from the enclosed mass plus 5% Gaussian dispersion, masses U[1,10]."""
    return generate("exponential-disk", DS_N, 100.0, 1.0, 10.0, 1.0, seed=seed)


def two_disks(seed: int = 3) -> Particles:
    """Two exponential disks of 250,000 (radius 33, scale 7) centred at
    (+-40, 0), approaching with bulk velocities -+2, each rotating."""
    return generate("two-exponential-disks", BH_N, 33.0, 1.0, 10.0, 1.0, seed=seed)


def batched_systems(count: int = 256, seed: int = 4):
    """The batched workload: `count` systems, N log-uniform in [1k, 64k],
    random-disk or exponential-disk (radius 100) chosen per system, masses
    U[1,10]; N < 8192 -> direct sum, else Barnes-Hut with theta = 0.7.
    Returns [(particles, algorithm, keyword parameters)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(64000))))
        kind = "random" if rng.uniform() < 0.5 else "exponential-disk"
        p = generate(kind, n, 100.0, 1.0, 10.0, 1.0, seed=1000 + k)
        algo = "direct-sum" if n < 8192 else "barnes-hut"
        out.append((p, algo, {"accuracy_parameter": 0.7} if algo == "barnes-hut" else {}))
    return out


@dataclass
class Request:
    at: float          # arrival time, seconds after the stream starts
    particles: Particles
    theta: float
    steps: int
    kind: str


def served_requests(seconds: float, rate: float = 50.0, seed: int = 7,
                    n_min: int = 2000, n_max: int = 2_000_000):
    """The served workload: Poisson arrivals at `rate`/s for `seconds`; N
    log-uniform in [n_min, n_max]; preset random, galaxy or galaxy-collision
    (size 100, masses U[1,10]); Barnes-Hut with theta in {0.5, 0.7, 1.0};
    steps uniform in {10..1000}.  Returns [Request]."""
    rng = np.random.default_rng(seed)
    kinds = ("random", "galaxy", "galaxy-collision")
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        kind = kinds[int(rng.integers(3))]
        theta = float((0.5, 0.7, 1.0)[int(rng.integers(3))])
        steps = int(rng.integers(10, 1001))
        p = generate(kind, n, 100.0, 1.0, 10.0, 1.0, seed=50_000 + len(out))
        out.append(Request(t, p, theta, steps, kind))
