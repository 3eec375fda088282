"""The allocation-free energy entry point (mfree_energy_noalloc): its
partial sums live in a fixed device array instead of memory the call
allocates.  Checked through the C ABI against mfree_energy on the same
bodies, on consecutive calls with a smaller input second, and its size
limit on the host."""
import ctypes as C

import numpy as np
import pytest

from mfree_iwf_amd import PRECISION_FP64, Solver, _capi, scenes
from mfree_iwf_amd._capi import Particles

FIELDS = ("m", "rho", "vx", "vy", "Sxx", "Sxy", "Syy")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    links (the test allocates its device inputs; the entry point does not)."""

    def __init__(self):
        _capi.lib()  # the library's runtime is loaded first
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _noalloc(ps, params, hip):
    view = Particles()
    for f in _capi.PARTICLE_FIELDS:
        setattr(view, f, C.cast(hip.put(getattr(ps, f)), _capi.dptr) if f in FIELDS else None)
    out = hip.put(np.full(3, np.nan))
    st = _capi.lib().mfree_energy_noalloc(0, None, C.byref(params), C.byref(view), ps.n, out)
    assert st == _capi.OK, _capi.lib().mfree_status_string(st).decode()
    e = hip.get(out, 3)
    hip.free()
    return e


def test_noalloc_rejects_inputs_beyond_its_scratch():
    # host-side check, before any device work
    p = Particles()
    for f in _capi.PARTICLE_FIELDS:
        setattr(p, f, C.cast(C.c_void_p(8), _capi.dptr))
    prm = scenes.rings_config1().params
    st = _capi.lib().mfree_energy_noalloc(0, None, C.byref(prm), C.byref(p), (1 << 24) + 1, C.c_void_p(8))
    assert st == _capi.ERR_SCRATCH
    assert "scratch" in _capi.lib().mfree_status_string(st).decode()


@pytest.mark.gpu
def test_noalloc_energy_equals_mfree_energy():
    hip = _Hip()
    # velocity jitter: density and stress change from the first step
    big = scenes.make_rings(241, shuffle=True, seed=1, jitter=2.0)             # 39k particles: 10 chunks
    small = scenes.make_rings(61, v0=40.0, E=3e7, shuffle=True, seed=2, jitter=2.0)  # fewer, other constants
    bodies = [big, small]
    print(f"[noalloc energy] n = {[b.particles.n for b in bodies]}", flush=True)
    with Solver(precision=PRECISION_FP64) as sv:
        for b in bodies:
            sv.add_body(b.particles, b.params)
        # before any step the storage is each body's idx order: the same
        # arrays in the same order, so the sums are bit-identical
        ref = [np.array(sv.energy(k)) for k in range(2)]
        got = [_noalloc(b.particles, b.params, hip) for b in bodies]  # two calls, the second smaller
        for r, g in zip(ref, got):
            assert np.array_equal(r, g), (r, g)
        assert ref[0][0] > 0 and ref[1][0] > 0
        # after steps (storage in cell order) against the state read back in
        # idx order: the same terms summed in another order
        sv.step(20)
        for k, b in enumerate(bodies):
            st = sv.state(k)
            e_ctx = np.array(sv.energy(k))
            e_na = _noalloc(st, b.params, hip)
            np.testing.assert_allclose(e_na, e_ctx, rtol=1e-12, atol=1e-300)
            assert e_na[1] > 0 and e_na[2] > 0
        # a larger call after a smaller one, then the smaller again: no state
        # carried over in the scratch
        again = [_noalloc(b.particles, b.params, hip) for b in (big, small, big)]
        assert np.array_equal(again[0], got[0]) and np.array_equal(again[1], got[1])
        assert np.array_equal(again[2], got[0])
