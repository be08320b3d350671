"""gsim_ds_step_noalloc: the direct-sum step that allocates nothing, called
through the C ABI on device arrays from the HIP runtime, against gsim_step on
a handle with the same one-sided launch shape (bit-exact: same kernels, same
slice count, same constants) and against the oracle's FP32 step (within the
1e-5 system-relative force tolerance).  Two calls in a row, the second with a
smaller input, share the static scratch."""
import ctypes as C
import os

import numpy as np
import pytest

import gsim
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for path in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = C.CDLL(path)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime not found")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipStreamSynchronize.argtypes = [C.c_void_p]
    lib.hipStreamDestroy.argtypes = [C.c_void_p]
    return lib


class DeviceSet:
    """pos, vel, mass in and pos, vel out, as device arrays of na particles."""

    def __init__(self, hip, p: gsim.Particles):
        self.hip = hip
        self.na = p.n
        self.bufs = {}
        for name, arr in (("pos", p.pos), ("vel", p.vel), ("mass", p.mass),
                          ("pos_out", p.pos), ("vel_out", p.vel)):
            a = np.ascontiguousarray(arr, np.float32)
            ptr = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr), a.nbytes) == 0
            self.bufs[name] = (ptr, a.nbytes)
            if not name.endswith("_out"):
                assert hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def addr(self, name: str) -> int:
        return self.bufs[name][0].value

    def read(self, name: str, shape) -> np.ndarray:
        out = np.empty(shape, np.float32)
        ptr, nbytes = self.bufs[name]
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, nbytes, D2H) == 0
        return out

    def free(self):
        for ptr, _ in self.bufs.values():
            self.hip.hipFree(ptr)


def _params(G, dt, soft):
    p = gsim.default_params()
    p.algorithm = gsim.DIRECT_SUM
    p.gravitational_const = G
    p.simulation_time = dt
    p.softening_len = soft
    return p


def _handle_step(pp: gsim.Particles, G, dt, soft):
    with gsim.Simulation(pp, "direct-sum", gravitational_const=G, simulation_time=dt,
                         softening_len=soft) as s:
        s.set_ds_shape(2, 0)  # the one-sided kernel with the default slice count
        s.step(1)
        return s.read()


@pytest.mark.parametrize("first,second", [((50_000, "exponential-disk", 1.0, 1e-3, 0.2),
                                           (3_000, "random", 0.5, 2e-3, 0.3))])
def test_noalloc_step_matches_gsim_step_twice(first, second):
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        for call, (n, kind, G, dt, soft) in enumerate((first, second)):
            print(f"noalloc call {call}: n={n} {kind} G={G} dt={dt} soft={soft}", flush=True)
            p = gsim.generate(kind, n, 100.0, 1.0, 10.0, seed=40 + call)
            pp = p.padded()
            dev = DeviceSet(hip, pp)
            try:
                gsim.ds_step_noalloc(_params(G, dt, soft), dev.addr("pos"), dev.addr("vel"),
                                     dev.addr("mass"), dev.addr("pos_out"),
                                     dev.addr("vel_out"), pp.n,
                                     stream.value if call == 0 else None)
                assert hip.hipStreamSynchronize(stream) == 0
                assert hip.hipDeviceSynchronize() == 0
                po = dev.read("pos_out", (pp.n, 2))
                vo = dev.read("vel_out", (pp.n, 2))
            finally:
                dev.free()
            ref = _handle_step(pp, G, dt, soft)
            assert np.array_equal(po, ref.pos), f"call {call}: positions differ from gsim_step"
            assert np.array_equal(vo, ref.vel), f"call {call}: velocities differ from gsim_step"
            # against the oracle: the velocity change is a * dt * G
            eps2 = H.sq32(soft)
            opo, ovo = O.ds_step(pp.pos, pp.vel, pp.mass, G, dt, eps2)
            dv, odv = vo - pp.vel, ovo - pp.vel
            assert H.system_rel_err(dv, odv) < 1e-5
            assert np.abs(po - opo).max() <= 1e-5 * np.abs(pp.pos).max()
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_rejects_bad_sizes_and_aliasing():
    hip = _hip()
    p = gsim.generate("random", 128, 100.0, 1.0, 10.0, seed=3).padded()
    dev = DeviceSet(hip, p)
    try:
        prm = _params(1.0, 1e-3, 0.2)
        args = [dev.addr(k) for k in ("pos", "vel", "mass", "pos_out", "vel_out")]
        with pytest.raises(gsim.GsimError) as e:
            gsim.ds_step_noalloc(prm, *args, 100)  # not a multiple of 64
        assert e.value.status == 1
        with pytest.raises(gsim.GsimError) as e:
            gsim.ds_step_noalloc(prm, *args, (32 << 20) // 8 + 64)  # beyond the scratch
        assert e.value.status == 6
        assert "scratch" in str(e.value)
        with pytest.raises(gsim.GsimError) as e:
            gsim.ds_step_noalloc(prm, args[0], args[1], args[2], args[0], args[4], 128)
        assert e.value.status == 1
        bh = _params(1.0, 1e-3, 0.2)
        bh.algorithm = gsim.BARNES_HUT
        with pytest.raises(gsim.GsimError) as e:
            gsim.ds_step_noalloc(bh, *args, 128)
        assert e.value.status == 1
    finally:
        dev.free()
