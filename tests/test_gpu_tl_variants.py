"""GPU parity of the co-rotated and potential-based Total-Lagrangian action
lists (tensile_corot, tensile_potential: src/benchmarks/tensile.cpp:137-361)
against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import _oracle as O
from mfree_iwf_amd import METHOD_TL_COROTATIONAL, METHOD_TL_POTENTIAL, Solver, scenes

pytestmark = pytest.mark.gpu

VARIANTS = [pytest.param(METHOD_TL_COROTATIONAL, O.TL_COROTATIONAL, id="corot"),
            pytest.param(METHOD_TL_POTENTIAL, O.TL_POTENTIAL, id="potential")]


def plate(dx=0.002, hdx=1.3, dt=None):
    return scenes.make_plate(dx=dx, L=0.1, W=0.04, r_hole=0.01, hdx=hdx, E=1e7, nu=0.4, rho0=1.0, v_bc=10.0, dt=dt)


def solver(method, pl, ps=None, bc=True):
    sv = Solver(method=method)
    sv.add_body(pl.particles if ps is None else ps, pl.params)
    if bc:
        sv.set_strip_bc(0, pl.left, pl.right, pl.v_bc)
    return sv


def oracle(variant, pl, ps=None, bc=True):
    ps = pl.particles if ps is None else ps
    t = O.TL(ps, float(ps.h[0]), search=1, contrib_only=1, variant=variant)
    assert t.rc == O.OK
    if bc:
        t.set_bc(pl.left, pl.right, pl.v_bc)
    return t


def perturbed(pl, seed):
    # positions moved off the reference configuration (H != 0, and a rotation
    # per particle for the co-rotated list), random velocities (viscosity)
    rng = np.random.default_rng(seed)
    ps = pl.particles.copy()
    n = ps.n
    ps.x += 0.05 * pl.dx * rng.uniform(-1, 1, n)
    ps.y += 0.05 * pl.dx * rng.uniform(-1, 1, n)
    ps.vx += rng.uniform(-1, 1, n)
    ps.vy += rng.uniform(-1, 1, n)
    ps.rho *= 1 + 1e-3 * rng.uniform(-1, 1, n)
    return ps


@pytest.mark.parametrize("method,variant", VARIANTS)
@pytest.mark.parametrize("seed", [0, 1])
def test_derivatives_vs_oracle(method, variant, seed):
    pl = plate()
    ps = perturbed(pl, seed)
    print(f"[tl variant deriv] method={method} n={ps.n} seed={seed}", flush=True)
    # lists and kernels from the configuration the body starts in, as the
    # drivers' set-up does (x = X there); derivatives at the perturbed state
    t = oracle(variant, pl)
    with solver(method, pl) as sv:
        sv.tl_precompute()
        sv.set_state(0, ps)
        d8, aux = sv.derivatives()
    rc, od8, oaux, _ = t.derivatives(ps, pl.params, time=0.0)
    assert rc == O.OK
    for k in range(8):
        assert O.rel_l2(d8[:, k], od8[:, k]) < 1e-12, k
    for k in range(16):
        assert O.rel_l2(aux[:, k], oaux[:, k]) < 1e-12, k
    assert np.all(d8[:, 4:] == 0.0)
    assert np.abs(oaux[:, 0]).max() > 1e-3  # a real deformation


@pytest.mark.parametrize("method,variant", VARIANTS)
def test_tension_driver_scene_vs_oracle(method, variant):
    # the reference main's tensile_corot(21) / tensile_potential(21): a unit
    # square pulled at 10 m/s, 300 steps
    pl = scenes.reference_tensile(21)
    print(f"[tl variant steps] method={method} n={pl.particles.n}", flush=True)
    t = oracle(variant, pl)
    with solver(method, pl) as sv:
        sv.step(300)
        g = sv.state(0)
        tg = sv.time(0)
    rc, o, to = t.step(pl.particles, pl.params, 300, threads=16)
    assert rc == O.OK
    ep = O.rel_l2(np.concatenate([g.x, g.y]), np.concatenate([o.x, o.y]))
    ev = O.rel_l2(np.concatenate([g.vx, g.vy]), np.concatenate([o.vx, o.vy]))
    print(f"  rel L2 vs oracle: pos {ep:.2e} vel {ev:.2e}", flush=True)
    assert ep <= 1e-10 and ev <= 1e-10
    # the interior moved (not only the strips)
    inner = np.setdiff1d(np.arange(pl.particles.n), np.concatenate([pl.left, pl.right]))
    assert np.abs(o.vx[inner]).max() > 1e-2
    # these lists integrate neither S nor rho
    assert np.array_equal(g.Sxx, pl.particles.Sxx) and np.array_equal(g.rho, pl.particles.rho)
    assert abs(tg - to) < 1e-18


@pytest.mark.parametrize("method,variant", VARIANTS)
def test_plate_with_hole_steps_vs_oracle(method, variant):
    # the potential list's forces are the gradient of its discrete potential
    # sum_i V_i psi(F_i) divided by a particle volume (src/potential.h:45-66:
    # Fe = -2 V F S enters without the neighbour's V): at this resolution
    # (V = 4e-6 m^2) the plate's CFL step diverges within 10 steps, in the
    # oracle as well; dt = 1e-9 s keeps it finite
    pl = plate() if method == METHOD_TL_COROTATIONAL else plate(dt=1e-9)
    print(f"[tl variant plate] method={method} n={pl.particles.n}", flush=True)
    t = oracle(variant, pl)
    with solver(method, pl) as sv:
        sv.step(100)
        g = sv.state(0)
    rc, o, _ = t.step(pl.particles, pl.params, 100, threads=16)
    assert rc == O.OK
    ep = O.rel_l2(np.concatenate([g.x, g.y]), np.concatenate([o.x, o.y]))
    ev = O.rel_l2(np.concatenate([g.vx, g.vy]), np.concatenate([o.vx, o.vy]))
    assert ep <= 1e-10 and ev <= 1e-10, (ep, ev)


def test_batched_bodies_equal_single_bodies():
    a, b = scenes.reference_tensile(15), scenes.reference_tensile(21)
    out = {}
    with Solver(method=METHOD_TL_COROTATIONAL) as sv:
        for k, s in enumerate((a, b)):
            sv.add_body(s.particles, s.params)
            sv.set_strip_bc(k, s.left, s.right, s.v_bc)
        sv.step(20)
        out["batch"] = [sv.state(0), sv.state(1)]
    for k, s in enumerate((a, b)):
        with solver(METHOD_TL_COROTATIONAL, s) as sv:
            sv.step(20)
            q = sv.state(0)
        for f in ("x", "y", "vx", "vy"):
            assert np.array_equal(getattr(q, f), getattr(out["batch"][k], f)), (k, f)
