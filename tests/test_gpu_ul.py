"""GPU parity of the Gray-Monaghan (UL) path against the CPU oracle, through
the C ABI."""
import numpy as np
import pytest

import _oracle as O
from _states import perturbed_rings
from mfree_iwf_amd import MfreeError, PRECISION_FP64, PRECISION_MIXED, Solver, _capi, scenes

pytestmark = pytest.mark.gpu


def solver_for(scene, precision=PRECISION_FP64, **kw):
    sv = Solver(precision=precision, **kw)
    sv.add_body(scene.particles, scene.params)
    return sv


def pos_stress_err(a, b):
    ep = O.rel_l2(np.concatenate([a.x, a.y]), np.concatenate([b.x, b.y]))
    es = O.rel_l2(np.concatenate([a.Sxx, a.Sxy, a.Syy]), np.concatenate([b.Sxx, b.Sxy, b.Syy]))
    return ep, es


@pytest.mark.parametrize("case", ["config1", "rings61_shuffled", "rings41_dense"])
def test_pairs_bit_exact(case):
    if case == "config1":
        s = scenes.rings_config1()
    elif case == "rings61_shuffled":
        s = perturbed_rings(nbox=61, seed=11, dpos=0.45, shuffle=True)
    else:
        s = perturbed_rings(nbox=41, seed=12, dpos=0.3, hdx=2.2)
    print(f"[pairs] {case} n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        row, col = sv.pairs(0)
    ps = s.particles
    rc, orow, ocol = O.verlet_pairs(ps.x, ps.y, ps.h, order=O.ORDER_STABLE)
    assert rc == O.OK
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    # and as sets against the reference's own emission order
    rc, rrow, rcol = O.verlet_pairs(ps.x, ps.y, ps.h, order=O.ORDER_REFERENCE)
    assert np.array_equal(row, rrow)
    for i in range(0, ps.n, max(1, ps.n // 500)):
        assert np.array_equal(np.sort(col[row[i]:row[i + 1]]), np.sort(rcol[rrow[i]:rrow[i + 1]]))


def test_pairs_config2_full_size():
    s = scenes.rings_config2()
    print(f"[pairs] config2 n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        row, col = sv.pairs(0)
    ps = s.particles
    rc, orow, ocol = O.verlet_pairs(ps.x, ps.y, ps.h, order=O.ORDER_STABLE, cap=ps.n * 70)
    assert rc == O.OK
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    cnt = np.diff(row.astype(np.int64))
    assert 55 < cnt.mean() < 65 and cnt.max() <= 64


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_derivatives_fp64_vs_oracle(seed):
    s = perturbed_rings(nbox=61, seed=seed, dv=5.0, dS=5e5, shuffle=bool(seed % 2), random_mass=seed == 2)
    print(f"[deriv64] seed={seed} n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        d8, aux = sv.derivatives()
    rc, od8, oaux = O.ul_derivatives(s.particles, s.params, order=O.ORDER_STABLE)
    assert rc == O.OK
    for k in range(8):
        assert O.rel_l2(d8[:, k], od8[:, k]) < 1e-12, k
    for k in range(13):
        assert O.rel_l2(aux[:, k], oaux[:, k]) < 1e-12, k
    assert np.array_equal(aux[:, 13], oaux[:, 13])  # reference neighbour counts


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_dense_derivatives_vs_oracle_and_xcd_chunks(precision):
    """h = 2.2 dx: about 60 neighbours inside the cutoff, so every lane fills
    its 32-entry neighbour buffer at least once and the collective and guard
    flushes of the pair pass both run; the result must not depend on how
    tiles are dealt to the XCDs either (one tile per chunk, the default
    64-tile chunks, one contiguous range per XCD)."""
    s = perturbed_rings(nbox=61, seed=21, dpos=0.3, dv=5.0, dS=5e5, hdx=2.2, shuffle=True)
    print(f"[dense] precision={precision} n={s.particles.n}", flush=True)
    out = []
    for chunk in (1, 0, 1 << 20):
        with solver_for(s, precision=precision, xcd_chunk=chunk) as sv:
            d8, _ = sv.derivatives()
            sv.step(2)
            out.append((d8, sv.state(0).state8()))
    for d8, st in out[1:]:
        assert np.array_equal(d8, out[0][0]) and np.array_equal(st, out[0][1])
    rc, od8, _ = O.ul_derivatives(s.particles, s.params, order=O.ORDER_STABLE)
    assert rc == O.OK
    tol = 1e-12 if precision == PRECISION_FP64 else 2e-4
    for k in range(8):
        assert O.rel_l2(out[0][0][:, k], od8[:, k]) < tol, k


def test_fp64_100_steps_config1():
    s = scenes.rings_config1()
    print(f"[fp64 100 steps] n={s.particles.n}", flush=True)
    with solver_for(s) as sv:
        sv.step(100)
        g = sv.state(0)
        t = sv.time(0)
    rc, o = O.ul_step(s.particles, s.params, 100, order=O.ORDER_STABLE)
    assert rc == O.OK
    ep, es = pos_stress_err(g, o)
    print(f"[fp64 100 steps] rel L2 pos {ep:.3e} stress {es:.3e}", flush=True)
    assert ep <= 1e-10 and es <= 1e-10
    assert abs(t - 100 * s.params.dt) < 1e-18
    assert np.abs(o.Sxx).max() > 1e3  # the rings are in contact: a real test


def test_fp64_perturbed_shuffled_steps():
    s = perturbed_rings(nbox=81, seed=21, dv=3.0, dS=1e5, shuffle=True, jitter=0.01)
    with solver_for(s) as sv:
        sv.step(40)
        g = sv.state(0)
    rc, o = O.ul_step(s.particles, s.params, 40, order=O.ORDER_STABLE)
    assert rc == O.OK
    ep, es = pos_stress_err(g, o)
    assert ep <= 1e-10 and es <= 1e-10


def test_graph_equals_eager_bitwise():
    s = perturbed_rings(nbox=61, seed=3)
    out = []
    for g in (True, False):
        with solver_for(s, use_graph=g) as sv:
            sv.step(7)
            out.append(sv.state(0).state8())
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_batched_bodies_equal_single_bodies():
    a = perturbed_rings(nbox=41, seed=4)
    b = perturbed_rings(nbox=61, seed=5, shuffle=True)
    with Solver() as sv:
        sv.add_body(a.particles, a.params)
        sv.add_body(b.particles, b.params)
        sv.step(5)
        ga, gb = sv.state(0), sv.state(1)
    for s, g in ((a, ga), (b, gb)):
        with solver_for(s) as sv:
            sv.step(5)
            e = sv.state(0)
        assert np.array_equal(g.state8().view(np.uint64), e.state8().view(np.uint64))


def test_set_state_roundtrip_and_restart():
    s = perturbed_rings(nbox=41, seed=6)
    with solver_for(s) as sv:
        sv.step(3)
        mid = sv.state(0)
        sv.step(3)
        end1 = sv.state(0)
        sv.set_state(0, mid)
        sv.step(3)
        end2 = sv.state(0)
    assert np.array_equal(end1.state8(), end2.state8())


def test_mixed_precision_error_vs_oracle():
    s = scenes.rings_config1()
    with solver_for(s, precision=PRECISION_MIXED) as sv:
        sv.step(100)
        g = sv.state(0)
    rc, o = O.ul_step(s.particles, s.params, 100, order=O.ORDER_STABLE)
    ep, es = pos_stress_err(g, o)
    print(f"[mixed 100 steps] rel L2 pos {ep:.3e} stress {es:.3e}", flush=True)
    assert ep <= 1e-5 and es <= 1e-3


def test_mixed_derivatives_close():
    s = perturbed_rings(nbox=61, seed=13, dv=5.0, dS=5e5)
    with solver_for(s, precision=PRECISION_MIXED) as sv:
        d8, aux = sv.derivatives()
    rc, od8, oaux = O.ul_derivatives(s.particles, s.params)
    for k in range(8):
        e = O.rel_l2(d8[:, k], od8[:, k])
        assert e < 2e-4, (k, e)


def test_error_neighbour_overflow():
    s = scenes.make_rings(41, ri=0.0, hdx=5.0)  # discs: ~300 neighbours per particle
    rc, _, _ = O.verlet_pairs(s.particles.x, s.particles.y, s.particles.h, cap=s.particles.n * 400)
    assert rc == O.ERR_NBH_OVERFLOW  # the reference asserts here (src/grid.h:167)
    with solver_for(s) as sv:
        with pytest.raises(MfreeError) as ei:
            sv.step(1)
    assert ei.value.status in (_capi.ERR_NBH_OVERFLOW, _capi.ERR_CELLS)


def test_error_cells_overflow():
    s = scenes.make_rings(41)
    ps = s.particles
    ps.x[0] = 50.0  # one particle far away: num_cell >= n (src/grid.h:114)
    with solver_for(s) as sv:
        with pytest.raises(MfreeError) as ei:
            sv.step(1)
    assert ei.value.status == _capi.ERR_CELLS


def test_nonfinite_detected():
    s = scenes.make_rings(41)
    s.particles.Sxx[5] = 1e308
    s.particles.Syy[5] = -1e308
    with solver_for(s) as sv:
        with pytest.raises(MfreeError) as ei:
            sv.step(1)
    assert ei.value.status == _capi.ERR_NONFINITE


@pytest.mark.parametrize("tc", [1, 5, 16, 32])
def test_mixed_tiles_match_untiled(tc):
    # the LDS-staged tile kernel and the per-particle kernel evaluate the same
    # FP32 sums (only the order of loads differs)
    s = perturbed_rings(nbox=81, seed=14, dv=5.0, dS=5e5, shuffle=True)
    with solver_for(s, precision=PRECISION_MIXED, tile_cells=tc) as sv:
        d8, _ = sv.derivatives(with_aux=False)
        assert sv.tile_cells == tc
    with solver_for(s, precision=PRECISION_MIXED, tile_cells=-1) as sv:
        e8, _ = sv.derivatives(with_aux=False)
        assert sv.tile_cells == 0
    rc, od8, _ = O.ul_derivatives(s.particles, s.params)
    for k in range(8):
        assert O.rel_l2(d8[:, k], e8[:, k]) < 1e-6, k
        assert O.rel_l2(d8[:, k], od8[:, k]) < 2e-4, k


@pytest.mark.parametrize("tc", [4, 32])
def test_mixed_tiles_dense(tc):
    # ~19 particles per cell, ~60 neighbours inside the cutoff: with 4-cell
    # tiles the candidates are staged in LDS and every lane's neighbour buffer
    # fills (collective and guard flushes); with 32-cell tiles they exceed the
    # LDS capacity and the tile is evaluated from global memory
    s = perturbed_rings(nbox=41, seed=12, dpos=0.3, hdx=2.2)
    with solver_for(s, precision=PRECISION_MIXED, tile_cells=tc) as sv:
        d8, _ = sv.derivatives(with_aux=False)
    rc, od8, _ = O.ul_derivatives(s.particles, s.params)
    for k in range(8):
        assert O.rel_l2(d8[:, k], od8[:, k]) < 2e-4, k


@pytest.mark.parametrize("tc", [-1, 16])
def test_mixed_larger_cells(tc):
    """Cells of 1.25 x 2 h_max (an autotuner candidate): the 3x3 block still
    holds every pair inside the cutoff, only with more candidates, so the
    derivatives agree with the oracle and with the reference-sized cells, and
    the Verlet-list helper still returns the reference's lists bit for bit."""
    s = perturbed_rings(nbox=61, seed=23, dv=5.0, dS=5e5, shuffle=True)
    with solver_for(s, precision=PRECISION_MIXED, tile_cells=tc, cell_scale=1.25) as sv:
        d8, _ = sv.derivatives(with_aux=False)  # the first compute call fixes the cell size
        assert sv.cell_scale == 1.25
        row, col = sv.pairs(0)
        sv.step(3)
        a = sv.state(0)
    with solver_for(s, precision=PRECISION_MIXED, tile_cells=-1) as sv:
        e8, _ = sv.derivatives(with_aux=False)
        assert sv.cell_scale == 1.0
        sv.step(3)
        b = sv.state(0)
    rc, od8, _ = O.ul_derivatives(s.particles, s.params)
    for k in range(8):
        assert O.rel_l2(d8[:, k], e8[:, k]) < 1e-6, k
        assert O.rel_l2(d8[:, k], od8[:, k]) < 2e-4, k
    ps = s.particles
    rc, orow, ocol = O.verlet_pairs(ps.x, ps.y, ps.h, order=O.ORDER_STABLE)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    ep, es = pos_stress_err(a, b)
    assert ep < 1e-9 and es < 1e-5


def test_mixed_autotuned_steps_vs_oracle():
    # per-stage binning (the autotuner chooses that pass's shape)
    s = perturbed_rings(nbox=81, seed=15, dv=2.0, dS=1e5, shuffle=True)
    with solver_for(s, precision=PRECISION_MIXED, neighbor_list=-1) as sv:
        print(f"[autotune] tile_cells={sv.tile_cells}", flush=True)
        sv.step(20)
        g = sv.state(0)
    rc, o = O.ul_step(s.particles, s.params, 20)
    ep, es = pos_stress_err(g, o)
    assert ep < 1e-6 and es < 1e-3


def test_batched_uneven_scenes_vs_oracle():
    # bodies of the batched workload (different sizes, speeds, moduli and time
    # steps) advanced together; every body matches its own oracle run
    specs = scenes.batch_specs(12, seed=3, nmin=500, nmax=6000)
    bodies = [scenes.make_rings(sp["nbox"], ri=sp["ri"], ro=sp["ro"], centre=sp["centre"], v0=sp["v0"], E=sp["E"])
              for sp in specs]
    print(f"[batch] {len(bodies)} bodies, {sum(b.particles.n for b in bodies)} particles", flush=True)
    with Solver() as sv:
        for b in bodies:
            sv.add_body(b.particles, b.params)
        sv.step(10)
        got = [sv.state(k) for k in range(len(bodies))]
        times = [sv.time(k) for k in range(len(bodies))]
    for b, g, t in zip(bodies, got, times):
        rc, o = O.ul_step(b.particles, b.params, 10)
        assert rc == O.OK
        ep, es = pos_stress_err(g, o)
        assert ep <= 1e-10 and es <= 1e-10
        assert abs(t - 10 * b.params.dt) < 1e-18
