"""Host-side checks of the C ABI that need no GPU: every declared symbol is
exported, scene generators, argument validation, and a loud failure of the
compute entry points without a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from mfree_iwf_amd import Solver, _capi, device_count, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mfree_iwf_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfree_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in _capi.SIGNATURES, f"{name} not bound in _capi"


def test_abi_version_and_status_strings():
    L = _capi.lib()
    assert L.mfree_abi_version() == 1
    assert L.mfree_status_string(_capi.ERR_CELLS).decode().startswith("cell grid")


def test_scene_counts():
    assert scenes.rings_config1().particles.n == 4408
    s2 = scenes.rings_config2()
    assert s2.particles.n == 1009204 and s2.nbox == 1213
    # shuffled, idx = position; every lattice point appears once
    s1 = scenes.make_rings(1213)
    assert np.array_equal(np.sort(s1.particles.x), np.sort(s2.particles.x))
    pl = scenes.plate_config3()
    assert 570000 < pl.particles.n < 585000
    assert len(pl.left) == len(pl.right) == 3 * 501


def test_scene_jitter_statistics():
    s = scenes.rings_config2()
    base = np.where(s.particles.x < 0, 59.0, -59.0)
    jx = s.particles.vx - base
    assert abs(jx.mean()) < 1e-4 and abs(jx.std() - 0.01) < 1e-4
    assert abs(s.particles.vy.std() - 0.01) < 1e-4


def test_cfl_time_step():
    s = scenes.rings_config1()
    c0 = scenes.sound_speed(1e7, 0.4, 1010.0)
    assert abs(s.params.dt - 0.3 * 1.3e-3 / (c0 + 59.0)) < 1e-18


def test_convex_hull_equals_the_oracle():
    # the library's host hull (the drivers' outline) against the oracle's,
    # which is pinned to the reference's own convex_hull.h
    import _oracle as O
    from mfree_iwf_amd import convex_hull
    rng = np.random.default_rng(2)
    g = np.arange(11, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="ij")
    for s in scenes.reference_rings_contact(41) + scenes.reference_rings_contact(21, "potential"):
        assert np.array_equal(convex_hull(s.particles.x, s.particles.y), O.convex_hull(s.particles.x, s.particles.y))
    for x, y in ((X.ravel(), Y.ravel()), (rng.uniform(size=300), rng.uniform(size=300))):
        assert np.array_equal(convex_hull(x, y), O.convex_hull(x, y))


def test_godunov_is_fp64_fused_only():
    # the Godunov pair sums exist in FP64 on the fused pipeline; other
    # combinations are refused at creation (no device needed)
    for kw in (dict(precision=_capi.PRECISION_MIXED), dict(pipeline=_capi.PIPELINE_REFERENCE)):
        with pytest.raises(_capi.MfreeError) as e:
            Solver(method=_capi.METHOD_UL_GODUNOV, **kw)
        assert e.value.status == _capi.ERR_UNSUPPORTED
    Solver(method=_capi.METHOD_UL_GODUNOV).close()


@pytest.mark.skipif(device_count() > 0, reason="checks the no-device path")
def test_compute_fails_loudly_without_device():
    s = scenes.make_rings(21)
    with Solver() as sv:
        sv.add_body(s.particles, s.params)
        with pytest.raises(_capi.MfreeError) as ei:
            sv.step(1)
    assert ei.value.status == _capi.ERR_HIP


def test_argument_validation():
    s = scenes.make_rings(21)
    with Solver() as sv:
        ps = s.particles.copy()
        ps.h[3] = -1.0
        with pytest.raises(_capi.MfreeError):
            sv.add_body(ps, s.params)
        bad = _capi.Params.from_buffer_copy(s.params)
        bad.dt = 0.0
        with pytest.raises(_capi.MfreeError):
            sv.add_body(s.particles, bad)


@pytest.mark.skipif(device_count() > 0, reason="checks the no-device path")
def test_cpp_driver_fails_loudly_without_device():
    exe = os.path.join(ROOT, "tests", "cpp", "reference_main")
    if not os.path.exists(exe):
        pytest.skip("C++ driver not built")
    r = subprocess.run([exe, "rings", "21", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "no HIP device" in r.stderr
