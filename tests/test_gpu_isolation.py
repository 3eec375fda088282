"""Per-body fault isolation in batched contexts (isolate_bodies): a body that
fails is frozen at the state it started its failing step with, its time
stops, and the other bodies step on exactly as they would alone.  Without
isolation the first failure fails the step (the reference asserts) and the
error names the body."""
import numpy as np
import pytest

from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, MfreeError, Solver, _capi, scenes

pytestmark = pytest.mark.gpu


def _good(seed):
    return scenes.make_rings(41, shuffle=True, seed=seed, jitter=0.3)


def _bad_late():
    # Mach ~20 impact of 1,684 particles: more than 200 neighbours (the
    # reference's MAX_NBH) after about 100 steps
    return scenes.make_rings(51, v0=1000.0)


def _solo(scene, nsteps, precision=PRECISION_FP64):
    with Solver(precision=precision) as sv:
        sv.add_body(scene.particles, scene.params)
        sv.step(nsteps)
        return sv.state(0).state8(), sv.time(0)


def _first_failure(scene, limit=200):
    """Steps a body alone, one step at a time, to its first failing step:
    (that step, its status, the state before it)."""
    with Solver() as sv:
        sv.add_body(scene.particles, scene.params)
        for k in range(1, limit + 1):
            before = sv.state(0).state8()
            try:
                sv.step(1)
            except MfreeError as e:
                return k, e.status, before
    raise AssertionError("the body did not fail")


def test_failed_body_is_frozen_and_the_others_step_on():
    a, b, c = _good(1), _bad_late(), _good(2)
    k_fail, status, before = _first_failure(b)
    print(f"[isolation] body 1 fails at step {k_fail} with status {status}", flush=True)
    assert status == _capi.ERR_NBH_OVERFLOW and 50 < k_fail < 150
    nsteps = 160
    with Solver(isolate_bodies=True) as sv:
        for s in (a, b, c):
            sv.add_body(s.particles, s.params)
        for _ in range(4):
            sv.step(nsteps // 4)  # every call succeeds
        got = [sv.state(k).state8() for k in range(3)]
        times = [sv.time(k) for k in range(3)]
        assert sv.failed_bodies() == {1: _capi.ERR_NBH_OVERFLOW}
        assert sv.body_status(0) == 0 and sv.body_status(2) == 0
    # the failed body: the state of its last good step, and that step's time
    assert np.array_equal(got[1], before)
    t = 0.0
    for _ in range(k_fail - 1):
        t = t + b.params.dt
    assert times[1] == t
    # the others: exactly their runs alone
    for k, s in ((0, a), (2, c)):
        alone, t_alone = _solo(s, nsteps)
        assert np.array_equal(got[k], alone), k
        assert times[k] == t_alone


def test_without_isolation_the_batch_fails_and_names_the_body():
    a, b = _good(3), _bad_late()
    with Solver() as sv:
        sv.add_body(a.particles, a.params)
        sv.add_body(b.particles, b.params)
        with pytest.raises(MfreeError) as ei:
            sv.step(160)
        assert ei.value.status == _capi.ERR_NBH_OVERFLOW
        assert "body 1" in str(ei.value)
        assert sv.body_status(1) == 0  # statuses are kept by isolating contexts only


def test_set_state_thaws_a_failed_body():
    a, b = _good(4), scenes.make_rings(41, ri=0.0, hdx=5.0)  # a disc with ~300 neighbours a particle
    with Solver(isolate_bodies=True) as sv:
        sv.add_body(a.particles, a.params)
        sv.add_body(b.particles, b.params)
        sv.step(3)
        assert sv.failed_bodies() == {1: _capi.ERR_NBH_OVERFLOW}
        assert np.array_equal(sv.state(1).state8(), b.particles.state8())
        assert sv.time(1) == 0.0
        # a state of the body's own size that steps: the disc with h = 1.3 dx
        ps = b.particles.copy()
        ps.h[:] = b.particles.h / 5.0 * 1.3
        sv.set_state(1, ps)
        assert sv.body_status(1) == 0
        sv.step(2)
        assert sv.body_status(1) == 0
        assert sv.time(1) == pytest.approx(2 * b.params.dt, rel=1e-12)


def test_mixed_list_path_freezes_a_list_overflow():
    # mixed precision (neighbour-list steps): a disc whose lists exceed the
    # list capacity (512 entries within 2 h + skin) at the first step: h = 8 dx
    # gives about 1,000
    a, b, c = _good(6), scenes.make_rings(41, ri=0.0, hdx=8.0), _good(7)
    with Solver(precision=PRECISION_MIXED, isolate_bodies=True) as sv:
        for s in (a, b, c):
            sv.add_body(s.particles, s.params)
        sv.step(10)
        assert sv.failed_bodies() == {1: _capi.ERR_NBH_OVERFLOW}
        got = [sv.state(k).state8() for k in range(3)]
    assert np.array_equal(got[1], b.particles.state8())
    for k, s in ((0, a), (2, c)):
        alone, _ = _solo(s, 10, PRECISION_MIXED)
        assert np.array_equal(got[k], alone), k


def test_isolation_options():
    from mfree_iwf_amd import METHOD_TL_RL
    with pytest.raises(MfreeError):
        Solver(method=METHOD_TL_RL, isolate_bodies=True)
    with pytest.raises(MfreeError):
        Solver(pipeline=_capi.PIPELINE_REFERENCE, isolate_bodies=True)
