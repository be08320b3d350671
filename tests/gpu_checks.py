"""GPU-side checks shared by the workload tests: one system stepped alone
through the C ABI, with chosen steps checked against the oracle evaluated on
the GPU's own input state of that step.

Per checked step:
  direct sum   forces within 1e-5 system-relative of the FP32 oracle, and per
               particle no less accurate against FP64 than the FP32 oracle
               itself (the GPU sums in another order); integrate bit-exact on
               the GPU's forces (SimShader.comp:68-69).
  Barnes-Hut   the tree bit-exact against the oracle's build, real-lane node
               visits and acceptances equal to the oracle's walk, forces as
               above (ForceShader.comp:111-141), integrate bit-exact
               (ForceShader.comp:145-146).
"""
from __future__ import annotations

import sys

import numpy as np

import gsim
import helpers as H
import oracle as O

FORCE_TOL = 1e-5


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def same(a: gsim.Particles, b: gsim.Particles) -> None:
    assert np.array_equal(a.pos, b.pos)
    assert np.array_equal(a.vel, b.vel)
    assert np.array_equal(a.mass, b.mass)


def _forces_ok(a, ref32, ref64, tol=FORCE_TOL):
    sysr = H.system_rel_err(a, ref32)
    assert sysr <= tol, f"system rel err {sysr:.3e}"
    eg = H.per_particle_rel(a, ref64)
    e32 = H.per_particle_rel(ref32, ref64)
    assert eg.max() <= max(1.5 * e32.max(), 1e-6), (eg.max(), e32.max())
    return sysr


def step_alone_checked(p: gsim.Particles, algo: str, kw: dict, steps: int,
                       check: tuple[int, ...]) -> tuple[gsim.Particles, list]:
    """Steps `p` (already padded to a multiple of 64, so the whole device
    state reads back) alone for `steps` steps; before each step index in
    `check` the state is read and that step is checked against the oracle.
    Returns the final state and one record per checked step."""
    assert p.n % 64 == 0, "pass padded particles: padding moves and shapes the tree"
    G = float(kw.get("gravitational_const", 1.0))
    dt = float(kw.get("simulation_time", 1e-3))
    soft = float(kw.get("softening_len", 0.2))
    theta = float(kw.get("accuracy_parameter", 1.0))
    speed = float(kw.get("simulation_speed", 1.0))
    W = int(kw.get("group_width", 64))
    eps2, theta2 = H.sq32(soft), H.sq32(theta)
    f = np.float32
    records = []
    s = gsim.Simulation(p, algo, **kw)
    try:
        s.set_debug(gsim.DEBUG_ACCEL)
        done = 0
        for k in sorted(set(check)):
            assert k < steps
            if k > done:
                s.step(k - done)
                done = k
            st = s.read()
            s.step(1)
            done += 1
            a = s.dump(gsim.DUMP_ACCEL)
            g = s.read()
            if algo == "direct-sum":
                ref32 = O.ds_accel_f32(st.pos, st.mass, eps2)
                ref64 = O.ds_accel_f64(st.pos, st.mass, eps2)
                err = _forces_ok(a, ref32, ref64)
                kv = f(dt) * f(G)
                nv = (st.vel + a * kv).astype(f)
                assert np.array_equal(g.vel, nv), f"step {k}: integrate (velocity)"
                np_ = (st.pos + ((st.vel + nv) * f(0.5)) * f(dt)).astype(f)
                assert np.array_equal(g.pos, np_), f"step {k}: integrate (position)"
                records.append({"step": k, "force_err": err})
            else:
                V, A = (int(x) for x in s.dump(gsim.DUMP_COUNTERS))
                t = s.tree()
                ot, nn, nin, pk, mk = O.bh_build(st.pos, st.mass)
                assert t["total"] == nn, f"step {k}: node count"
                for key in ("sorted_src", "counts", "radii", "node_pos", "node_mass"):
                    assert np.array_equal(t[key], ot[key]), f"step {k}: tree {key}"
                oa, oV, oA, _, _, oa64 = O.bh_traverse(t, pk, mk, eps2, theta2, W=W, f64=True)
                assert (V, A) == (oV, oA), f"step {k}: visits/accepts {(V, A)} vs {(oV, oA)}"
                inside = t["sorted_src"][t["sorted_src"] < len(mk)]
                err = _forces_ok(a[inside], oa[inside], oa64[inside])
                dtb = f(dt) * f(speed)
                kv, kp = f(G) * dtb, f(0.5) * dtb
                nv = (st.vel[inside] + a[inside] * kv).astype(f)
                assert np.array_equal(g.vel[inside], nv), f"step {k}: integrate (velocity)"
                assert np.array_equal(g.pos[inside],
                                      (pk[inside] + (st.vel[inside] + nv) * kp).astype(f))
                # killed particles (outside [-500, 500)^2): parked, massless
                assert np.array_equal(g.mass, mk), f"step {k}: kill masses"
                out = np.setdiff1d(np.arange(len(mk)), inside)
                assert np.array_equal(g.pos[out], pk[out]), f"step {k}: kill positions"
                records.append({"step": k, "force_err": err, "visits": V, "accepts": A})
            say(f"    step {k}: force err {err:.2e}" +
                (f", V {records[-1]['visits']} A {records[-1]['accepts']}"
                 if algo != "direct-sum" else ""))
        if steps > done:
            s.step(steps - done)
        return s.read(), records
    finally:
        s.close()
