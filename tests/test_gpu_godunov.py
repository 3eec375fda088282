"""GPU parity of the Godunov (Parshikov) UL method -- the rings_godunov action
list, src/benchmarks/rings.cpp:151-158 with src/godunov.h:41-189 -- against
the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import _oracle as O
from _states import perturbed_rings
from mfree_iwf_amd import METHOD_UL_GODUNOV, Solver, scenes

pytestmark = pytest.mark.gpu

# the constants of rings_godunov (src/benchmarks/rings.cpp:101-161)
GD = dict(centre=0.041, hdx=1.7, rho0=1.0, v0=180.0, dt=1e-7, alpha=0.0, beta=0.0, eta=0.0, art_eps=0.0,
          corr_exp=0.0, xsph_eps=0.0)
FIELDS = ("x", "y", "vx", "vy", "rho", "Sxx", "Sxy", "Syy")


def gd_solver(*scene_list, **kw):
    sv = Solver(method=METHOD_UL_GODUNOV, **kw)
    for s in scene_list:
        sv.add_body(s.particles, s.params)
    return sv


@pytest.mark.parametrize("seed", [0, 1])
def test_derivatives_vs_oracle(seed):
    # every field perturbed, order shuffled: pressures, stresses and
    # velocities differ between every pair
    s = perturbed_rings(nbox=41, seed=seed, dv=5.0, drho=0.01, dS=2e5, shuffle=True, **GD)
    print(f"[godunov derivatives] seed={seed} n={s.particles.n}", flush=True)
    with gd_solver(s) as sv:
        d8, aux = sv.derivatives()
    rc, od8, oaux = O.ul_derivatives(s.particles, s.params, method=O.METHOD_GODUNOV)
    assert rc == O.OK
    for k in range(8):
        e = O.rel_l2(d8[:, k], od8[:, k])
        assert e < 1e-12, (k, e)
    for k in (0, 1, 9, 10, 11, 12):  # quadrature weight, p, grad v
        assert O.rel_l2(aux[:, k], oaux[:, k]) < 1e-12, k
    assert np.array_equal(aux[:, 13], oaux[:, 13])  # neighbour counts
    assert np.all(aux[:, 2:9] == 0.0)                # no artificial stress, no stress divergence
    assert np.array_equal(d8[:, 0], s.particles.vx)  # advection only


def test_100_steps_vs_oracle():
    # the reference main's call, rings_godunov(80): the rings touch after
    # about 15 steps
    s = scenes.reference_rings_godunov(80)
    print(f"[godunov 100 steps] n={s.particles.n}", flush=True)
    with gd_solver(s) as sv:
        sv.step(100)
        g = sv.state(0)
        t = sv.time(0)
    rc, o = O.ul_step(s.particles, s.params, 100, threads=16, method=O.METHOD_GODUNOV)
    assert rc == O.OK
    ep = O.rel_l2(np.concatenate([g.x, g.y]), np.concatenate([o.x, o.y]))
    es = O.rel_l2(np.concatenate([g.Sxx, g.Sxy, g.Syy]), np.concatenate([o.Sxx, o.Sxy, o.Syy]))
    ev = O.rel_l2(np.concatenate([g.vx, g.vy]), np.concatenate([o.vx, o.vy]))
    er = O.rel_l2(g.rho, o.rho)
    print(f"  rel L2 vs oracle: pos {ep:.2e} vel {ev:.2e} stress {es:.2e} rho {er:.2e}", flush=True)
    assert ep < 1e-10 and es < 1e-10 and ev < 1e-10 and er < 1e-10
    assert np.max(o.rho) > 1.01  # the collision has started
    assert abs(t - 100 * 1e-7) < 1e-18


def test_batched_bodies_equal_single_bodies():
    a = perturbed_rings(nbox=31, seed=5, dv=5.0, shuffle=True, **GD)
    b = perturbed_rings(nbox=41, seed=6, dv=5.0, **GD)
    with gd_solver(a, b) as sv:
        sv.step(5)
        ga, gb = sv.state(0), sv.state(1)
    for s, g in ((a, ga), (b, gb)):
        with gd_solver(s) as sv:
            sv.step(5)
            q = sv.state(0)
        for f in FIELDS:
            assert np.array_equal(getattr(q, f), getattr(g, f)), f


def test_graph_equals_eager_bitwise():
    s = perturbed_rings(nbox=41, seed=7, dv=5.0, shuffle=True, **GD)
    out = []
    for graph in (True, False):
        with gd_solver(s, use_graph=graph) as sv:
            sv.step(6)
            out.append(sv.state(0))
    for f in FIELDS:
        assert np.array_equal(getattr(out[0], f), getattr(out[1], f)), f
