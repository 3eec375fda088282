"""GPU checks of the learned stepper against its FP64 oracle (oracle/mpnn.py)
and the plain solver, through the C ABI: graph builder, MFMA forward and
backward (gradient check), the learned step with and without the network,
the judge / rollback, training, and the ablation against the base update."""
import numpy as np
import pytest

import _oracle as O
from _states import perturbed_rings
from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver, scenes
from mfree_iwf_amd import surrogate as sg
from oracle import mpnn as om

pytestmark = pytest.mark.gpu

# BF16 operands (8-bit mantissa) in every dense layer and in the aggregated
# messages, FP32 accumulation: the network's outputs and gradients agree with
# the FP64 oracle to about 1e-2 relative.
Y_TOL = 3e-2
GRAD_TOL = 6e-2


def lattice_graph(n, seed, hdx=1.3):
    """Jittered lattice; positions rounded to float32 (what the GPU sees)."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    ij = np.array([(i, j) for i in range(side) for j in range(side)][:n], dtype=np.float64)
    p = (ij + 0.2 * rng.standard_normal(ij.shape)).astype(np.float32).astype(np.float64)
    h = np.full(n, hdx)
    pos3 = np.stack([p[:, 0], p[:, 1], h], axis=1)
    return pos3, om.radius_graph(p[:, 0], p[:, 1], h)


def sorted_rows(row, col, e3):
    out = []
    for i in range(len(row) - 1):
        a, b = int(row[i]), int(row[i + 1])
        o = np.argsort(col[a:b], kind="stable")
        out.append((col[a:b][o], e3[a:b][o]))
    return out


def test_graph_matches_oracle():
    s = perturbed_rings(nbox=41, seed=3, dpos=0.3)
    ps = s.particles
    print(f"[surrogate graph] n={ps.n}", flush=True)
    with Solver(precision=PRECISION_FP64) as sv, sg.Surrogate() as su:
        sv.add_body(ps, s.params)
        row, col, e3 = su.graph(sv)
    orow, ocol, oe3, _ = om.radius_graph(ps.x, ps.y, ps.h)
    # the edge set is decided in FP64 (bit-exact); the weights are evaluated in
    # FP32 from FP32 positions (a few 1e-6 relative to the weight scale)
    assert np.array_equal(row.astype(np.int64), orow)
    for (c, e), (oc, oe) in zip(sorted_rows(row, col, e3), sorted_rows(orow, ocol, oe3)):
        assert np.array_equal(c, oc)
        assert np.allclose(e, oe, rtol=1e-4, atol=2e-5)


@pytest.mark.parametrize("case", ["perturbed", "lattice_ties"])
def test_graph_from_the_mixed_pair_pass_equals_graph_builder(case):
    """Mixed contexts of uniform-h bodies take the learned step's edges from
    the D1 pair pass; they equal the FP64 graph builder's edge for edge, in
    order, also on a lattice whose neighbours sit exactly at r = 2h."""
    if case == "perturbed":
        s = perturbed_rings(nbox=41, seed=6, dpos=0.3)
    else:
        s = scenes.make_rings(41, hdx=1.5)  # 2h = 3 dx: lattice distance (3, 0)
    ps = s.particles
    print(f"[surrogate fused graph] {case} n={ps.n}", flush=True)
    out = {}
    for prec in (PRECISION_FP64, PRECISION_MIXED):
        with Solver(precision=prec) as sv, sg.Surrogate() as su:
            sv.add_body(ps, s.params)
            out[prec] = su.graph(sv)
    for a, b in zip(out[PRECISION_FP64], out[PRECISION_MIXED]):
        assert np.array_equal(a, b)
    orow, ocol, _, _ = om.radius_graph(ps.x, ps.y, ps.h)
    row, col, _ = out[PRECISION_MIXED]
    assert np.array_equal(row.astype(np.int64), orow)
    for k in range(ps.n):
        assert np.array_equal(np.sort(col[row[k]:row[k + 1]]), np.sort(ocol[orow[k]:orow[k + 1]]))


def test_graph_multibody_matches_oracle():
    bodies = [perturbed_rings(nbox=nb, seed=20 + k, dpos=0.3) for k, nb in enumerate((21, 41, 31))]
    print(f"[surrogate graph] bodies n={[b.particles.n for b in bodies]}", flush=True)
    with Solver(precision=PRECISION_MIXED) as sv, sg.Surrogate() as su:
        for b in bodies:
            sv.add_body(b.particles, b.params)
        sv.step(3)
        states = [sv.state(k) for k in range(3)]
        row, col, e3 = su.graph(sv)
    off = 0
    for ps in states:
        orow, ocol, oe3, _ = om.radius_graph(ps.x, ps.y, ps.h)
        r = row[off:off + ps.n + 1].astype(np.int64) - int(row[off])
        assert np.array_equal(r, orow)
        a, b = int(row[off]), int(row[off + ps.n])
        got = sorted_rows(r, col[a:b].astype(np.int64) - off, e3[a:b])
        for (c, e), (oc, oe) in zip(got, sorted_rows(orow, ocol, oe3)):
            assert np.array_equal(c, oc)
            assert np.allclose(e, oe, rtol=1e-4, atol=2e-5)
        off += ps.n


@pytest.mark.parametrize("n,seed", [(300, 0), (1000, 1)])
def test_mpnn_forward_backward_matches_oracle(n, seed):
    pos3, (row, col, e3, _) = lattice_graph(n, seed)
    rng = np.random.default_rng(seed + 10)
    X = rng.standard_normal((n, om.NF))
    T = rng.standard_normal((n, om.NOUT))
    w = om.init_params(seed=seed + 20).astype(np.float32)
    print(f"[mpnn fwd/bwd] n={n} edges={len(col)}", flush=True)
    y, loss, g = sg.forward_backward(row, col, pos3, X, T, w)
    # oracle on the same float32 weights (FP64 arithmetic)
    oy, oloss, og = om.loss_and_grad(row, col, e3, X, T, w.astype(np.float64))
    assert O.rel_l2(y, oy) < Y_TOL
    assert loss == pytest.approx(oloss, rel=Y_TOL)
    off = 0
    for (r, c) in om.layout():
        k = r * c
        assert O.rel_l2(g[off:off + k], og[off:off + k]) < GRAD_TOL, (r, c)
        off += k
    cos = float(np.dot(g, og) / (np.linalg.norm(g) * np.linalg.norm(og)))
    assert cos > 0.998


def ragged_graph(n, seed):
    """Jittered lattice with per-node smoothing lengths: isolated nodes,
    degrees up to the 64-edge limit, a partial last tile of 16 nodes."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    ij = np.array([(i, j) for i in range(side) for j in range(side)][:n], dtype=np.float64)
    p = (ij + 0.2 * rng.standard_normal(ij.shape)).astype(np.float32).astype(np.float64)
    h = rng.uniform(0.25, 2.0, n)
    h[(ij[:, 0] < 8) & (ij[:, 1] < 8)] = 2.2  # a block of dense rows
    p[-5:] = [(side + 50.0 + 10.0 * k, 0.0) for k in range(5)]  # isolated nodes
    h = h.astype(np.float32).astype(np.float64)
    g = om.radius_graph(p[:, 0], p[:, 1], h)
    return np.stack([p[:, 0], p[:, 1], h], axis=1), g


@pytest.mark.parametrize("n,seed", [(300, 0), (1000, 1), (4099, 2)])
def test_mpnn_inference_forward_matches_oracle(n, seed):
    """The fused layer kernels (aggregation in the MFMA operand layout, decoder
    in the last layer's epilogue) against the oracle and the training forward."""
    pos3, (row, col, e3, _) = ragged_graph(n, seed)
    deg = np.diff(row)
    assert deg.max() <= sg.MAX_DEG and (deg == 0).any() and deg.max() > 40
    rng = np.random.default_rng(seed + 10)
    X = rng.standard_normal((n, om.NF))
    w = om.init_params(seed=seed + 20).astype(np.float32)
    print(f"[mpnn inference fwd] n={n} edges={len(col)} max deg={deg.max()}", flush=True)
    y, _, _ = sg.forward_backward(row, col, pos3, X, None, w, train=False)
    yt, _, _ = sg.forward_backward(row, col, pos3, X, np.zeros((n, om.NOUT)), w)
    oy, _ = om.forward(row, col, e3, X, w.astype(np.float64))
    assert np.isfinite(y).all()
    assert O.rel_l2(y, oy) < Y_TOL
    # same sums; the training forward keeps an FP32 residual stream
    assert O.rel_l2(y, yt) < 1e-2


def _ctx(scene, precision=PRECISION_FP64):
    sv = Solver(precision=precision)
    sv.add_body(scene.particles, scene.params)
    return sv


def _scales(scene):
    p = scene.params
    K = scenes.elastic_K(p.E, p.nu)
    c0 = np.sqrt(K / p.rho0)
    return 4 * p.dt, c0, K, p.rho0


def _oracle_base(scene, ps, order=4):
    """The base step of state ps: one RK step (order 4: RK4, 3: Kutta) over
    tau = 4 dt of the oracle's UL derivatives (stable order).  Returns
    (start state, new state, stage rates)."""
    tau, _, _, _ = _scales(scene)

    def deriv(s8):
        q = ps.copy()
        for k, f in enumerate(scenes.STATE_FIELDS):
            setattr(q, f, np.ascontiguousarray(s8[:, k]))
        rc, d8, _ = O.ul_derivatives(q, scene.params)
        assert rc == O.OK
        return d8

    s8 = ps.state8()
    new, ks = om.rk_stages(s8, deriv, tau, order)
    return s8, new, ks


@pytest.mark.parametrize("order", [3, 4])
def test_base_step_matches_oracle(order):
    s = perturbed_rings(nbox=41, seed=5, dpos=0.2)
    with _ctx(s) as sv, sg.Surrogate(base_order=order) as su:
        su.record_velocity(sv)
        su.step(sv, use_net=False)
        got = sv.state(0).state8()
        t = sv.time(0)
    s8, base, _ = _oracle_base(s, s.particles, order)
    for k in range(8):
        # the increments, which carry the physics (the state itself is
        # dominated by the positions and the rest density)
        assert O.rel_l2(got[:, k] - s8[:, k], base[:, k] - s8[:, k]) < 1e-10, scenes.STATE_FIELDS[k]
    assert t == pytest.approx(4 * s.params.dt, rel=1e-12)


def test_base_step_mixed_close_to_fp64():
    """The mixed base step (FP32 compensated pair sums, FP64 stage updates)
    against the FP64 one."""
    s = perturbed_rings(nbox=41, seed=8, dpos=0.2)
    out = []
    for prec in (PRECISION_FP64, PRECISION_MIXED):
        with _ctx(s, prec) as sv, sg.Surrogate() as su:
            su.record_velocity(sv)
            su.step(sv, use_net=False)
            out.append(sv.state(0).state8())
    base = s.particles.state8()
    for k in range(8):
        # increments agree to the mixed mode's FP32 pair sums
        assert O.rel_l2(out[1][:, k] - base[:, k], out[0][:, k] - base[:, k]) < 1e-4, scenes.STATE_FIELDS[k]


@pytest.mark.parametrize("order", [3, 4])
def test_learned_step_matches_oracle(order):
    """Network step on random weights: the correction added to the base
    update agrees with the FP64 oracle's features -> forward -> correction
    -> conservation projection."""
    s = perturbed_rings(nbox=41, seed=6, dpos=0.2)
    ps = s.particles
    tau, c0, K, rho0 = _scales(s)
    s8, base, ks = _oracle_base(s, ps, order)
    row, col, e3, wsum = om.radius_graph(ps.x, ps.y, ps.h)
    feats = om.features(s8, ks, s8[:, 2:4], ps.h, wsum, tau, c0, K, rho0)
    fm, fs = feats.mean(0), feats.std(0)
    fs[fs < 1e-6 * np.abs(fm) + 1e-30] = 1.0  # constant features (the Courant number)
    norm = np.concatenate([fm, fs, np.zeros(om.NOUT), np.full(om.NOUT, 1e-3)]).astype(np.float32)
    w = om.init_params(seed=9).astype(np.float32)
    X = (feats - norm[:om.NF]) / norm[om.NF:2 * om.NF]
    Y, _ = om.forward(row, col, e3, X, w.astype(np.float64))
    want = om.conserve(om.correction(Y, norm.astype(np.float64), ps.h, c0, K, rho0), ps.m)
    with _ctx(s) as sv, sg.Surrogate(base_order=order) as su:
        su.set_params(w, norm)
        su.record_velocity(sv)
        su.step(sv, use_net=True)
        got = sv.state(0).state8() - base
    for k in range(8):
        assert O.rel_l2(got[:, k], want[:, k]) < Y_TOL, scenes.STATE_FIELDS[k]


def test_judge_always_rejecting_equals_plain_solver():
    s = scenes.rings_config1()
    with _ctx(s) as a, _ctx(s) as b, sg.Surrogate() as su:
        su.set_params(norm=np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.ones(8)]))
        st = su.rollout(a, 4 + 2 * 32 + 3, group=8, threshold=-1.0)
        b.step(4 + 2 * 32 + 3)
        assert st["groups"] == 2 and st["fallbacks"] == 2 and st["learned_steps"] == 0
        assert st["plain_steps"] == 4 + 2 * 32 + 3
        sa, sb = a.state(0).state8(), b.state(0).state8()
        assert np.array_equal(sa, sb)
        assert a.time(0) == pytest.approx(b.time(0), rel=1e-12)


def test_judge_rejects_a_bad_network_at_the_default_threshold():
    # a seeded network whose corrections are scaled to the state itself (target
    # std 1 in units of h, c0, rho0, K) moves particles by about h per step:
    # every group is rejected -- by the judge or as a numerical failure -- and
    # the rollout equals the plain solver bit for bit
    s = scenes.rings_config1()
    nsteps = 4 + 3 * 8 * 4 + 3
    with _ctx(s, PRECISION_MIXED) as a, _ctx(s, PRECISION_MIXED) as b, sg.Surrogate(seed=11) as su:
        su.set_params(norm=np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.ones(8)]))
        st = su.rollout(a, nsteps, group=8)  # default threshold
        print(f"[judge bad network] {st}", flush=True)
        b.step(nsteps)
        assert st["groups"] == 3 and st["fallbacks"] == 3 and st["learned_steps"] == 0
        assert st["plain_steps"] == nsteps
        # a body whose positions blew up takes no quadratic work before it fails
        # (its binning fails at once; no cell sort, no all-pairs scan)
        assert st["ms_learned"] < 5000.0
        assert np.array_equal(a.state(0).state8(), b.state(0).state8())


def test_rebind_to_a_new_context_of_the_same_size():
    """A learned stepper that outlives its context rebinds to the next one
    (contexts are told apart by id, not by address)."""
    s1 = perturbed_rings(nbox=41, seed=3, dpos=0.2)
    s2 = perturbed_rings(nbox=41, seed=4, dpos=0.2)
    assert s1.particles.n == s2.particles.n
    norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.full(8, 1e-3)])
    with sg.Surrogate(seed=5) as su:
        su.set_params(norm=norm)
        for sc in (s1, s2):
            with _ctx(sc, PRECISION_MIXED) as sv:
                su.step(sv, use_net=True)
                got = sv.state(0).state8()
            with _ctx(sc, PRECISION_MIXED) as sv, sg.Surrogate(seed=5) as fresh:
                fresh.set_params(norm=norm)
                fresh.step(sv, use_net=True)
                want = sv.state(0).state8()
            assert np.array_equal(got, want)


def test_rollout_accepting_counts_steps():
    s = scenes.rings_config1()
    with _ctx(s, PRECISION_MIXED) as a, sg.Surrogate() as su:
        su.set_params(norm=np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.full(8, 1e-9)]))
        st = su.rollout(a, 4 + 3 * 64 + 5, group=16, threshold=np.inf)
        assert st["groups"] == 3 and st["fallbacks"] == 0
        # kept learned steps: 15 per group (the judged one's result is the
        # plain steps'); every one of the 201 steps is counted once
        assert st["learned_steps"] == 45 and st["plain_steps"] == 4 + 3 * 4 + 5
        assert 4 * st["learned_steps"] + st["plain_steps"] == 4 + 3 * 64 + 5
        # the network's output is switched off: the judged step is its base step
        assert 0.0 < st["max_deviation"] < 1e-2
        assert st["max_ratio_stress"] <= 1.1
        t = a.time(0)
    assert t == pytest.approx((4 + 3 * 64 + 5) * s.params.dt, rel=1e-9)


def test_batched_rollout_equals_separate_rollouts():
    """A learned rollout of 3 bodies in one context equals the rollouts of
    each body alone, bit for bit (per-node features and network rows, per-body
    correction means, per-body pair passes).  The pair-pass shape is fixed in
    both (the autotuner times the passes, and its choice for three bodies may
    differ from that for one, which changes the order of the sums)."""
    bodies = scenes.scenes_from_specs(scenes.subsonic_batch_specs(nscenes=3, seed=5, nmin=1500, nmax=3000))
    w = om.init_params(seed=4).astype(np.float32)
    norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.full(8, 1e-3)])
    nsteps = 4 + 2 * 8 * 4 + 4
    fixed = dict(tile_cells=-1, cell_scale=1.0)
    with Solver(precision=PRECISION_MIXED, **fixed) as sv, sg.Surrogate(seed=2) as su:
        for b in bodies:
            sv.add_body(b.particles, b.params)
        su.set_params(w, norm)
        st = su.rollout(sv, nsteps, group=8, threshold=np.inf)
        assert st["fallbacks"] == 0 and st["learned_steps"] == 2 * 7  # the judged step of each group is replaced
        assert 4 * st["learned_steps"] + st["plain_steps"] == nsteps
        together = [sv.state(k).state8() for k in range(3)]
    for k, b in enumerate(bodies):
        with Solver(precision=PRECISION_MIXED, **fixed) as sv, sg.Surrogate(seed=2) as su:
            sv.add_body(b.particles, b.params)
            su.set_params(w, norm)
            su.rollout(sv, nsteps, group=8, threshold=np.inf)
            alone = sv.state(0).state8()
        assert np.array_equal(alone, together[k]), k


def _batch(nscenes, seed, nmin, nmax):
    return scenes.scenes_from_specs(scenes.subsonic_batch_specs(nscenes=nscenes, seed=seed, nmin=nmin, nmax=nmax))


def _twins(scene, k):
    return [_ctx(scene, PRECISION_MIXED) for _ in range(k)]


# median gain over the velocity and stress errors the network must reach on
# held-out scenes: Kutta's step leaves a larger, (tau A)^4-dominated error
# for the network to remove than RK4 over the same 4 dt does
ABLATION_GAIN = {3: 1.5, 4: 1.2}


@pytest.mark.parametrize("order", [4, 3])
def test_training_and_ablation(order):
    """Trains on FP64 rollouts of a small batch, then compares one learned
    step with and without the network against 4 RK4 steps on held-out
    scenes (mixed precision, as in a rollout)."""
    train = _batch(24, seed=1, nmin=2000, nmax=6000)
    print(f"[surrogate train] order {order} bodies={len(train)} n={sum(t.particles.n for t in train)}", flush=True)
    with Solver(precision=PRECISION_FP64) as sv, \
            sg.Surrogate(seed=3, max_batch_nodes=40000, base_order=order) as su:
        for sc in train:
            sv.add_body(sc.particles, sc.params)
        su.collect(sv, ncapture=24, gap=4)
        print(f"[surrogate train] usable snapshots {su.snapshots} of {24 * 24}", flush=True)
        assert 0.5 * 24 * 24 < su.snapshots <= 24 * 24
        loss = su.train(1000, lr=2e-3)
        print(f"[surrogate train] loss first {loss[:10].mean():.4f} last {loss[-50:].mean():.4f}", flush=True)
        assert np.all(np.isfinite(loss))
        assert loss[-50:].mean() < 0.7 * loss[:10].mean()
        gains = []
        for sc in _batch(3, seed=77, nmin=3000, nmax=8000):
            a, b, c = _twins(sc, 3)
            with a, b, c:
                for x in (a, b, c):
                    x.step(150)
                su.record_velocity(a)
                su.record_velocity(b)
                for x in (a, b, c):
                    x.step(4)
                su.step(a, use_net=True)
                su.step(b, use_net=False)
                c.step(4)
                ta, tb, tc = a.state(0).state8(), b.state(0).state8(), c.state(0).state8()
            d_net = np.linalg.norm(ta[:, 2:8] - tc[:, 2:8], axis=0)
            d_base = np.linalg.norm(tb[:, 2:8] - tc[:, 2:8], axis=0)
            gains.append(d_base / np.maximum(d_net, 1e-300))
            print(f"[ablation] n={sc.particles.n} base/net error ratio (vx vy rho Sxx Sxy Syy): "
                  + " ".join(f"{g:.2f}" for g in gains[-1]), flush=True)
        gains = np.array(gains)
        # the network reduces the error of the velocity and stress update
        assert np.median(gains[:, [0, 1, 3, 4, 5]]) > ABLATION_GAIN[order]
