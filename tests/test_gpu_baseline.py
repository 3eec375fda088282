"""GPU parity of the reference-structured FP64 pipeline (the "plain FP64
GPU" baseline: materialised lists, one kernel per action) against the
oracle and the fused FP64 path."""
import numpy as np
import pytest

import _oracle as O
from _states import perturbed_rings
from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, MfreeError, Solver, _capi, scenes

pytestmark = pytest.mark.gpu


def solver_for(scene, pipeline=_capi.PIPELINE_REFERENCE, **kw):
    sv = Solver(precision=PRECISION_FP64, pipeline=pipeline, **kw)
    sv.add_body(scene.particles, scene.params)
    return sv


@pytest.mark.parametrize("seed", [0, 1])
def test_reference_pipeline_derivatives_vs_oracle(seed):
    s = perturbed_rings(nbox=61, seed=seed, dv=5.0, dS=5e5, shuffle=bool(seed % 2), random_mass=seed == 1)
    print(f"[ref pipeline deriv] seed={seed} n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        d8, _ = sv.derivatives()
    rc, od8, _ = O.ul_derivatives(s.particles, s.params, order=O.ORDER_STABLE)
    assert rc == O.OK
    for k in range(8):
        assert O.rel_l2(d8[:, k], od8[:, k]) < 1e-12, k


def test_reference_pipeline_100_steps_vs_oracle_and_fused():
    s = scenes.rings_config1()
    print(f"[ref pipeline 100 steps] n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        sv.step(100)
        g = sv.state(0)
        t = sv.time(0)
    with solver_for(s, pipeline=_capi.PIPELINE_FUSED) as sv:
        sv.step(100)
        f = sv.state(0)
    rc, o = O.ul_step(s.particles, s.params, 100, order=O.ORDER_STABLE)
    assert rc == O.OK
    for a in (g, f):
        assert O.rel_l2(np.concatenate([a.x, a.y]), np.concatenate([o.x, o.y])) <= 1e-10
        assert O.rel_l2(np.concatenate([a.Sxx, a.Sxy, a.Syy]), np.concatenate([o.Sxx, o.Sxy, o.Syy])) <= 1e-10
    # same sums in the same order: the two GPU pipelines agree to rounding
    assert O.rel_l2(g.state8(), f.state8()) <= 1e-13
    assert abs(t - 100 * s.params.dt) < 1e-18


def test_reference_pipeline_options():
    with pytest.raises(MfreeError):
        Solver(precision=PRECISION_MIXED, pipeline=_capi.PIPELINE_REFERENCE)
    # neighbour overflow is reported like the fused path's (discs, ~300 neighbours)
    d = scenes.make_rings(41, ri=0.0, hdx=5.0)
    with solver_for(d) as sv:
        with pytest.raises(MfreeError) as e:
            sv.step(1)
    assert e.value.status in (_capi.ERR_NBH_OVERFLOW, _capi.ERR_CELLS)
