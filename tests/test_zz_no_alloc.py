"""The allocation-free energy entry point (sors_energy_noalloc) against
sors_energy, through the C ABI.

Device buffers for its inputs and output come from the HIP runtime the
library itself links (libamdhip64.so.7); the entry point allocates nothing
and keeps its reduction partials in the library's static device scratch.
Two calls in a row, the second on a smaller mesh, must each match
sors_energy bit for bit: the same kernels sum in the same order.
"""
import ctypes as C

import numpy as np
import pytest

from sors_amd import mesh, native

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs to own device buffers."""

    def __init__(self):
        native.lib()  # loads libamdhip64.so.7 as the library's dependency
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def to_device(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.check(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def empty(self, count):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), 8 * count), "hipMalloc")
        return p

    def to_host(self, p, count):
        out = np.empty(count, np.float64)
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.h.hipMemcpy(out.ctypes.data, p, 8 * count, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self.check(self.h.hipFree(p), "hipFree")


def _state(sc, seed):
    rng = np.random.default_rng(seed)
    q = sc.X.reshape(-1) + 1e-4 * rng.standard_normal(3 * sc.X.shape[0])
    return q


def _noalloc(r, hip, q, dt, stream=None):
    qd, out = hip.to_device(q), hip.empty(1)
    try:
        rc = native.lib().sors_energy_noalloc(r.h, qd, float(dt), out, stream)
        r._check(rc)
        return float(hip.to_host(out, 1)[0])
    finally:
        hip.free(qd)
        hip.free(out)


def test_noalloc_energy_matches_sors_energy_two_calls():
    hip = _Hip()
    big = mesh.beam(nx=40, ny=8, nz=8, steps=1)      # 15,360 tets: 60 reduction blocks
    small = mesh.beam(nx=6, ny=2, nz=2, steps=1)     # 144 tets: one block
    results = []
    for sc, seed in ((big, 1), (small, 2)):
        print(f"noalloc energy: {sc.n_elems} tets, seed {seed}", flush=True)
        r = native.Rollout.from_scene(sc)
        q = _state(sc, seed)
        for dt in (sc.dt, 0.0):
            e_ref = r.energy(q, dt)
            e_na = _noalloc(r, hip, q, dt)
            assert np.isfinite(e_ref)
            assert e_na == e_ref, (sc.n_elems, dt, e_na, e_ref)
            results.append(e_na)
        r.close()
    assert results[0] != results[2]


def test_noalloc_energy_after_steps_and_with_contact():
    """After steps (qprev, vprev moved) and with contact terms, on a small
    cube: the contact sum is part of the total, as in sors_energy."""
    hip = _Hip()
    sc = mesh.beam(nx=8, ny=2, nz=2, steps=2)
    r = native.Rollout.from_scene(sc)
    q = sc.X.reshape(-1).copy()
    for _ in range(2):
        q, _ = r.step(q, sc.dt)
    z0 = float(sc.X[:, 2].min())
    r.set_contact(planes=[((0.0, 0.0, z0 + 1e-4), (0.0, 0.0, 1.0))], stiffness=1e5)
    qt = q + 1e-5
    e_ref = r.energy(qt, sc.dt)
    e_na = _noalloc(r, hip, qt, sc.dt)
    assert e_na == e_ref
    r.close()


def test_noalloc_energy_rejects_null_arguments():
    hip = _Hip()
    sc = mesh.beam(nx=4, ny=2, nz=2, steps=1)
    r = native.Rollout.from_scene(sc)
    out = hip.empty(1)
    assert native.lib().sors_energy_noalloc(r.h, None, 1e-3, out, None) == native.ERR_ARG
    assert native.lib().sors_energy_noalloc(None, out, 1e-3, out, None) == native.ERR_ARG
    hip.free(out)
    r.close()
