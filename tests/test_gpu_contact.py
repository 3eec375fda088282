"""GPU parity of the penalty contact between bodies (src/contact.h:36-80,
src/boundary.h:47-73, the main loop src/mfree_iwf.cpp:84-97) against the CPU
oracle, through the C ABI: outlines, contact forces and the two-ring contact
drivers stepped with the main loop's cadence (step, outline update, every
ordered pair of bodies)."""
import numpy as np
import pytest

import _oracle as O
from _contact import KIND, OracleBodies
from mfree_iwf_amd import METHOD_UL_MONAGHAN, PRECISION_FP64, PRECISION_MIXED, Solver, convex_hull, scenes

pytestmark = pytest.mark.gpu


def two_ring_solver(bodies, method, precision=PRECISION_FP64):
    sv = Solver(method=method, precision=precision)
    hulls = []
    for k, b in enumerate(bodies):
        sv.add_body(b.particles, b.params)
        hulls.append(convex_hull(b.particles.x, b.particles.y))
        sv.set_boundary(k, hulls[-1])
    return sv, hulls


def test_outlines_and_forces_after_reordered_steps():
    # UL storage is re-sorted every step: the outline follows the particles
    # by idx; forces of pushed-together rings equal the oracle's bit for bit
    bodies = scenes.reference_rings_contact(41)
    for b in bodies:  # overlap the rings by 3 mm
        b.particles.x += -0.003 if b.particles.x[0] > 0 else 0.003
    sv, hulls = two_ring_solver(bodies, METHOD_UL_MONAGHAN)
    with sv:
        sv.step(3)
        sv.update_boundaries()
        st = [sv.state(k) for k in range(2)]
        for k in range(2):
            pts = O.boundary_update(hulls[k], st[k].x, st[k].y)
            np.testing.assert_array_equal(sv.boundary(k), pts)
        sv.contact_all()
        pts = [sv.boundary(k) for k in range(2)]
        for j in range(2):
            fx, fy = sv.contact_forces(j)
            ox, oy = O.contact(pts[1 - j], st[j].x, st[j].y, st[j].m, bodies[j].params.dt)
            np.testing.assert_array_equal(fx, ox)
            np.testing.assert_array_equal(fy, oy)
            assert np.count_nonzero(ox) > 10


@pytest.mark.parametrize("kind,nbox,steps", [("monaghan", 41, 30), ("godunov", 41, 250), ("corot", 41, 250)])
def test_contact_drivers_main_loop_vs_oracle(kind, nbox, steps):
    # at nbox = 41 the outlines start 6 mm apart at 360 m/s relative speed:
    # in contact after about 17 us (17 steps of 1e-6 s, 170 of 1e-7 s)
    bodies = scenes.reference_rings_contact(nbox, kind)
    print(f"[contact loop] {kind} n={bodies[0].particles.n} x2 steps={steps}", flush=True)
    orc = OracleBodies(bodies, kind)
    sv, _ = two_ring_solver(bodies, KIND[kind][0])
    touched = 0
    with sv:
        for _ in range(steps):
            sv.step(1)
            sv.update_boundaries()
            sv.contact_all()
            orc.step()
            touched += int(np.count_nonzero(orc.fc[0][0]) > 0)
        g = [sv.state(k) for k in range(2)]
        f = [sv.contact_forces(k) for k in range(2)]
    assert touched > 5  # the contact forces acted
    for k in range(2):
        o = orc.ps[k]
        ep = O.rel_l2(np.concatenate([g[k].x, g[k].y]), np.concatenate([o.x, o.y]))
        ev = O.rel_l2(np.concatenate([g[k].vx, g[k].vy]), np.concatenate([o.vx, o.vy]))
        ef = O.rel_l2(np.concatenate(f[k]), np.concatenate(orc.fc[k]))
        print(f"  body {k}: rel L2 pos {ep:.2e} vel {ev:.2e} forces {ef:.2e}", flush=True)
        assert ep <= 1e-10 and ev <= 1e-10 and ef <= 1e-8


def test_contact_between_contexts_equals_one_context():
    # the C++ adapter keeps one context per body: contact across contexts
    bodies = scenes.reference_rings_contact(31)
    for b in bodies:
        b.particles.x += -0.003 if b.particles.x[0] > 0 else 0.003
    one, hulls = two_ring_solver(bodies, METHOD_UL_MONAGHAN)
    with one:
        for _ in range(4):
            one.step(1)
            one.update_boundaries()
            one.contact_all()
        ref = [one.state(k) for k in range(2)], [one.contact_forces(k) for k in range(2)]
    sv = [Solver(), Solver()]
    try:
        for k in range(2):
            sv[k].add_body(bodies[k].particles, bodies[k].params)
            sv[k].set_boundary(0, hulls[k])
        for _ in range(4):
            for k in range(2):
                sv[k].step(1)
                sv[k].update_boundaries()
            sv[1].contact(0, 0, master=sv[0])
            sv[0].contact(0, 0, master=sv[1])
        for k in range(2):
            s = sv[k].state(0)
            np.testing.assert_array_equal(s.x, ref[0][k].x)
            np.testing.assert_array_equal(s.vx, ref[0][k].vx)
            fx, fy = sv[k].contact_forces(0)
            np.testing.assert_array_equal(fx, ref[1][k][0])
    finally:
        for s in sv:
            s.close()


def test_mixed_precision_contact_loop():
    # mixed contexts step on the neighbour-list path; the forces enter the
    # FP64 epilogue: close to the FP64 loop
    bodies = scenes.reference_rings_contact(41)
    out = {}
    for prec in (PRECISION_FP64, PRECISION_MIXED):
        sv, _ = two_ring_solver(bodies, METHOD_UL_MONAGHAN, precision=prec)
        with sv:
            for _ in range(30):
                sv.step(1)
                sv.update_boundaries()
                sv.contact_all()
            out[prec] = [sv.state(k) for k in range(2)]
    for k in range(2):
        a, b = out[PRECISION_MIXED][k], out[PRECISION_FP64][k]
        assert O.rel_l2(np.concatenate([a.x, a.y]), np.concatenate([b.x, b.y])) < 1e-6
        assert O.rel_l2(np.concatenate([a.vx, a.vy]), np.concatenate([b.vx, b.vy])) < 1e-3
