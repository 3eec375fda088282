"""Benchmark of the SPH particle-interaction hot path on one MI355X.

Workload: the ring collision at dx = 6.6e-5 m (nbox = 1213, 1,009,204
particles), idx order shuffled from seed 0, N(0, 0.01 m/s) velocity jitter,
E = 1e7 Pa, nu = 0.4, rho0 = 1010 kg/m^3, h = 1.3 dx, alpha = beta = 1,
artificial stress eps = 0.3, RK4 with a CFL-0.3 time step.

A step is one RK4 step of every particle (4 derivative evaluations,
re-binning included).  Modes:
  fp64   all arithmetic FP64 (the reference's arithmetic)
  mixed  FP32 pair sums with compensated accumulation, FP64 per-particle
         algebra and RK4 state
`value` is the mixed-mode particle-steps/s; every mode's rate, the mixed
mode's rel-L2 error against the FP64 mode after the accuracy horizon, the
dominant kernel's roofline and the CPU baseline are in the same JSON line.

    python bench.py [--steps K] [--warmup W] [--accuracy-steps A]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M particles: FP64 vs mixed vs surrogate at ≤1e-3 rel L2 error"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one stage launch of the pair kernel, per particle:
# read state8 + m, h (80 B) and write 8 derivatives (64 B).
ALG_BYTES_PER_PARTICLE_STAGE = 144.0
# algorithmic flops of one stage per particle: 128 per active pair (kernel 29,
# stress divergence 40, velocity gradient 14, viscosity 34, XSPH 11) times the
# 20.85 active pairs per particle of the 1M ring workload, plus 100 for the
# per-particle algebra and the RK update
N_FLOP_PER_PARTICLE_STAGE = 20.85 * 128 + 100
FP32_VALU_PEAK_TFLOPS = 157.3
FP64_VALU_PEAK_TFLOPS = 78.6
ALG_BYTES_PER_PARTICLE_STEP = 704.0


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def run_mode(scene, precision, steps, warmup, log, pipeline=0):
    from mfree_iwf_amd import Solver
    n = scene.particles.n
    with Solver(precision=precision, use_graph=True, pipeline=pipeline) as sv:
        sv.add_body(scene.particles, scene.params)
        t0 = time.perf_counter()
        sv.step(max(warmup, 1))
        log(f"  warmup {warmup} steps: {time.perf_counter() - t0:.2f} s (pair tile cells: {sv.tile_cells})")
        tile_cells = sv.tile_cells
        # timed region: HIP events around every pair-kernel launch on the
        # context's stream (no graph while they are recorded)
        sv.profile(True)
        t0 = time.perf_counter()
        sv.step(steps)  # returns after a stream synchronize
        wall = time.perf_counter() - t0
        prof = sv.profile_read()
        sv.profile(False)
        # the same number of steps replayed from the captured HIP graph
        t0 = time.perf_counter()
        sv.step(steps)
        wall_graph = time.perf_counter() - t0
        state = sv.state(0)
    ms = prof["step_ms"] / max(prof["steps"], 1)
    pair_ms = prof["pair_ms"] / max(prof["pair_launches"], 1)
    return dict(
        particle_steps_per_s=n * steps / (prof["step_ms"] / 1e3),
        particle_steps_per_s_graph=n * steps / wall_graph,
        ms_per_step=ms,
        wall_s=wall,
        pair_kernel_ms=pair_ms,
        pair_launches=prof["pair_launches"],
        tile_cells=tile_cells,
        state=state,
    )


def errors(a, b):
    """rel-L2 of state a against the FP64 reference state b."""
    return dict(
        rel_l2_pos=rel_l2(np.concatenate([a.x, a.y]), np.concatenate([b.x, b.y])),
        rel_l2_vel=rel_l2(np.concatenate([a.vx, a.vy]), np.concatenate([b.vx, b.vy])),
        rel_l2_stress=rel_l2(np.concatenate([a.Sxx, a.Sxy, a.Syy]), np.concatenate([b.Sxx, b.Sxy, b.Syy])),
        rel_l2_rho=rel_l2(a.rho, b.rho),
    )


# The oracle's end state of the workload after the accuracy horizon (idx
# order), when given with --oracle-state: every mode's error is then also
# reported against it.
ORACLE_STATE = None


def load_oracle_state(path):
    """x, y, vx, vy, rho, Sxx, Sxy, Syy arrays of an npz file written from the
    oracle's state (numpy arrays only, no pickles)."""
    from types import SimpleNamespace
    with np.load(path, allow_pickle=False) as z:
        return SimpleNamespace(**{k: np.array(z[k]) for k in ("x", "y", "vx", "vy", "rho", "Sxx", "Sxy", "Syy")})


def vs_oracle(state, res, key="error_vs_oracle"):
    if ORACLE_STATE is not None:
        res[key] = errors(state, ORACLE_STATE)


def accuracy(scene, nsteps, log):
    """Mixed vs FP64 GPU after nsteps RK4 steps from the same state.
    Returns (summary, FP64 end state)."""
    from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver
    out, energy = {}, {}
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        with Solver(precision=prec) as sv:
            sv.add_body(scene.particles, scene.params)
            e0 = sum(sv.energy(0))
            t0 = time.perf_counter()
            sv.step(nsteps)
            out[name] = sv.state(0)
            energy[name] = sum(sv.energy(0)) / e0
            log(f"  accuracy run {name}: {nsteps} steps in {time.perf_counter() - t0:.2f} s, "
                f"energy E/E0 = {energy[name]:.9f}")
    a, b = out["mixed"], out["fp64"]
    res = dict(steps=nsteps, **errors(a, b),
               max_abs_stress_pa=float(np.abs(np.concatenate([b.Sxx, b.Sxy, b.Syy])).max()),
               energy_ratio=energy)
    if ORACLE_STATE is not None:
        res["fp64_vs_oracle"] = errors(b, ORACLE_STATE)
        res["mixed_vs_oracle"] = errors(a, ORACLE_STATE)
        log(f"  vs oracle after {nsteps} steps: fp64 {res['fp64_vs_oracle']} mixed {res['mixed_vs_oracle']}")
    return res, b


def bench_surrogate(scene, args, ref_state, log):
    """Learned stepper: training data from the batched workload's generator
    (subsonic scenes, seed 0), training on the GPU, then the judged rollout
    of the full workload over the accuracy horizon, timed end to end."""
    from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver, Surrogate, scenes
    t0 = time.perf_counter()
    train = scenes.scenes_from_specs(scenes.subsonic_batch_specs(args.sur_scenes, seed=0))
    ntrain = sum(s.particles.n for s in train)
    log(f"surrogate: {len(train)} training bodies, {ntrain} particles (setup {time.perf_counter() - t0:.1f} s)")
    res = {}
    res["base_order"] = args.sur_order
    with Surrogate(seed=0, max_batch_nodes=args.sur_batch, base_order=args.sur_order) as su:
        if args.sur_iters == 0:
            # kernel timing only: the seeded network, unit feature normalisation
            su.set_params(norm=np.concatenate([np.zeros(16), np.ones(16), np.zeros(8), np.full(8, 1e-6)]))
            return _rollout(su, scene, args, None, res, log)
        with Solver(precision=PRECISION_FP64 if args.sur_collect == "fp64" else PRECISION_MIXED) as gen:
            for s in train:
                gen.add_body(s.particles, s.params)
            t0 = time.perf_counter()
            su.collect(gen, ncapture=args.sur_captures, gap=args.sur_gap)
            res["collect_s"] = time.perf_counter() - t0
        res["collect_precision"] = args.sur_collect
        res["snapshots"] = su.snapshots
        res["dataset_nodes"] = su.dataset_nodes
        log(f"  collected {su.snapshots} usable snapshots ({su.dataset_nodes} nodes) in {res['collect_s']:.1f} s")
        t0 = time.perf_counter()
        loss = su.train(args.sur_iters, lr=args.sur_lr)
        res["train_s"] = time.perf_counter() - t0
        res["train_iters"] = args.sur_iters
        res["loss_first"] = float(np.mean(loss[:20]))
        res["loss_last"] = float(np.mean(loss[-100:]))
        log(f"  trained {args.sur_iters} iterations in {res['train_s']:.1f} s, loss {res['loss_first']:.4f} -> "
            f"{res['loss_last']:.4f}")
        res = _rollout(su, scene, args, ref_state, res, log)
        if ref_state is not None and "error_vs_fp64" in res:
            res["base_step_only"] = _ablation(su, scene, args, ref_state, log)
        return res


def _ablation(su, scene, args, ref_state, log):
    """The learned stepper without its network: the same warm-up and the same
    4-dt base steps (no network kernels, no judge), timed end to end like the
    rollout, with its errors at the accuracy horizon."""
    from mfree_iwf_amd import PRECISION_MIXED, Solver
    n = scene.particles.n
    nsteps = args.accuracy_steps
    with Solver(precision=PRECISION_MIXED) as sv:
        sv.add_body(scene.particles, scene.params)
        sv.step(0)  # commit outside the timed region
        t0 = time.perf_counter()
        sv.step(4)
        nl = (nsteps - 4) // 4
        for _ in range(nl):
            su.step(sv, use_net=False)
        if nsteps - 4 - 4 * nl:
            sv.step(nsteps - 4 - 4 * nl)
        wall = time.perf_counter() - t0
        st = sv.state(0)
    err = errors(st, ref_state)
    res = dict(particle_steps_per_s=n * nsteps / wall, rollout_s=wall, steps=nsteps, base_steps=nl,
               error_vs_fp64=err)
    log(f"  base step alone: {nsteps} steps in {wall:.2f} s ({res['particle_steps_per_s']:.4e} particle-steps/s), "
        f"vs fp64: {err}")
    vs_oracle(st, res)
    return res


def _rollout(su, scene, args, ref_state, res, log):
    """The judged rollout of the full workload, timed end to end (the energy
    check only for a trained network)."""
    from mfree_iwf_amd import PRECISION_MIXED, Solver
    n = scene.particles.n
    nsteps = args.sur_steps if args.sur_steps > 0 else args.accuracy_steps
    trained = args.sur_iters > 0
    with Solver(precision=PRECISION_MIXED) as sv:
        sv.add_body(scene.particles, scene.params)
        sv.step(0)  # commit (autotuner) outside the timed region
        e0 = sum(sv.energy(0)) if trained else None
        t0 = time.perf_counter()
        st = su.rollout(sv, nsteps, group=args.sur_group, threshold=args.sur_threshold)
        wall = time.perf_counter() - t0
        state = sv.state(0)
        if trained:
            res["energy_ratio"] = sum(sv.energy(0)) / e0
    res.update(particle_steps_per_s=n * nsteps / wall, rollout_s=wall, steps=nsteps, judge=st)
    log(f"  rollout {nsteps} steps in {wall:.2f} s: {res['particle_steps_per_s']:.4e} particle-steps/s, "
        f"judge {st}")
    if ref_state is not None and nsteps == args.accuracy_steps:
        res["error_vs_fp64"] = errors(state, ref_state)
        log(f"  surrogate vs fp64 after {nsteps} steps: {res['error_vs_fp64']}")
        vs_oracle(state, res)
    return res


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its logical core count."""
    name = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(scene, steps, log):
    """The CPU oracle on one host thread, bounded sample: the reference's
    algorithm and action order in FP64 (per-stage std::sort in the reference's
    order, Verlet lists, one loop per action), restated over SoA arrays and
    built with g++ -O2 -- not the reference's own AoS build, so it runs faster
    than the reference would."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    t0 = time.perf_counter()
    rc, _ = O.ul_step(scene.particles, scene.params, steps, order=O.ORDER_REFERENCE, threads=1)
    dt = time.perf_counter() - t0
    if rc != O.OK:
        raise RuntimeError(f"oracle failed with status {rc}")
    log(f"  cpu baseline: {steps} steps in {dt:.1f} s")
    model, ncpu = cpu_model()
    return dict(value=scene.particles.n * steps / dt, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{scene.particles.n} particles (the full workload), {steps} RK4 steps; SoA oracle "
                       f"(g++ -O2) in the reference's sort order on 1 thread",
                cpu_model=model, host_logical_cpus=ncpu, seconds=dt)


def validate_fp64(scene, nsteps, log):
    """FP64 GPU mode vs the oracle (stable order, all host threads) after
    nsteps RK4 steps of the full workload: the 1e-10 rel-L2 gate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from mfree_iwf_amd import PRECISION_FP64, Solver
    t0 = time.perf_counter()
    with Solver(precision=PRECISION_FP64) as sv:
        sv.add_body(scene.particles, scene.params)
        sv.step(nsteps)
        g = sv.state(0)
    tg = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    rc, o = O.ul_step(scene.particles, scene.params, nsteps, order=O.ORDER_STABLE, threads=threads)
    to = time.perf_counter() - t0
    if rc != O.OK:
        raise RuntimeError(f"oracle failed with status {rc}")
    res = dict(steps=nsteps, particles=scene.particles.n, oracle_threads=threads, gpu_s=tg, oracle_s=to,
               rel_l2_pos=rel_l2(np.concatenate([g.x, g.y]), np.concatenate([o.x, o.y])),
               rel_l2_stress=rel_l2(np.concatenate([g.Sxx, g.Sxy, g.Syy]), np.concatenate([o.Sxx, o.Sxy, o.Syy])),
               rel_l2_vel=rel_l2(np.concatenate([g.vx, g.vy]), np.concatenate([o.vx, o.vy])),
               max_abs_stress_pa=float(np.abs(np.concatenate([o.Sxx, o.Sxy, o.Syy])).max()))
    log(f"fp64 validation: {res}")
    return res


def bench_batch(steps, warmup, nscenes, log, accuracy_steps=500):
    """Batched uneven bodies: nscenes ring collisions drawn from seed 0, all
    advanced by one launch per kernel per RK stage (HIP graph), bodies
    isolated: a body that fails (more than 200 neighbours, as the reference
    asserts, after the fastest impacts) is frozen and the batch steps on.
    Each precision's context is timed over its first steps, then continues to
    the configuration's horizon; mixed vs FP64 per scene at the horizon."""
    from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver, scenes
    t0 = time.perf_counter()
    sc = scenes.batch_config4(nscenes, 0)
    ntot = sum(s.particles.n for s in sc)
    log(f"batch: {nscenes} scenes, {ntot} particles (setup {time.perf_counter() - t0:.1f} s)")
    out, states, status = {}, {}, {}
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        with Solver(precision=prec, isolate_bodies=True) as sv:
            for s in sc:
                sv.add_body(s.particles, s.params)
            w = max(warmup, 1)
            sv.step(w)
            sv.profile(True)
            sv.step(steps)
            prof = sv.profile_read()
            sv.profile(False)
            t0 = time.perf_counter()
            sv.step(steps)
            wall_graph = time.perf_counter() - t0
            tc = sv.tile_cells
            r = dict(particle_steps_per_s=ntot * steps / (prof["step_ms"] / 1e3),
                     particle_steps_per_s_graph=ntot * steps / wall_graph,
                     ms_per_step=prof["step_ms"] / steps,
                     pair_kernel_ms=prof["pair_ms"] / max(prof["pair_launches"], 1), tile_cells=tc)
            done = w + 2 * steps
            if accuracy_steps > done:
                t0 = time.perf_counter()
                sv.step(accuracy_steps - done)
                wall = time.perf_counter() - t0
                r["horizon_particle_steps_per_s_graph"] = ntot * (accuracy_steps - done) / wall
                log(f"  {name}: steps {done}..{accuracy_steps} in {wall:.2f} s")
            states[name] = [sv.state(k) for k in range(len(sc))]
            status[name] = [sv.body_status(k) for k in range(len(sc))]
        out[name] = r
        log(f"  {name}: {r['particle_steps_per_s']:.4e} particle-steps/s ({r['ms_per_step']:.3f} ms/step), "
            f"graph {r['particle_steps_per_s_graph']:.4e}")
    if accuracy_steps > 0:
        ok = [k for k in range(len(sc)) if status["fp64"][k] == 0 and status["mixed"][k] == 0]
        per = [errors(states["mixed"][k], states["fp64"][k]) for k in ok]

        def tally(v):
            t = {}
            for x in v:
                if x:
                    t[str(x)] = t.get(str(x), 0) + 1
            return t
        acc = dict(steps=accuracy_steps, scenes=len(sc), ok_in_both=len(ok),
                   failed_fp64=tally(status["fp64"]), failed_mixed=tally(status["mixed"]),
                   within_1e3=sum(1 for e in per if max(e.values()) <= 1e-3),
                   within_1e3_pos_stress=sum(1 for e in per if e["rel_l2_pos"] <= 1e-3
                                             and e["rel_l2_stress"] <= 1e-3),
                   failed_scenes_fp64={k: s for k, s in enumerate(status["fp64"]) if s},
                   failed_scenes_mixed={k: s for k, s in enumerate(status["mixed"]) if s})
        if per:
            acc.update({f"max_{k}": max(e[k] for e in per) for k in per[0]})
            acc.update({f"median_{k}": float(np.median([e[k] for e in per])) for k in per[0]})
        out["accuracy"] = acc
        log(f"  mixed vs fp64 after {accuracy_steps} steps (per scene): {acc}")
    return ntot, out


def bench_plate(steps, warmup, log, accuracy_steps=2000):
    """Total-Lagrangian corrected SPH: the tension plate with a hole
    (0.1 x 0.04 m, r = 0.01 m, dx = 8e-5 m), strip velocity BCs."""
    from mfree_iwf_amd import METHOD_TL_RL, PRECISION_FP64, PRECISION_MIXED, Solver, scenes
    t0 = time.perf_counter()
    pl = scenes.plate_config3()
    n = pl.particles.n
    log(f"plate: {n} particles (setup {time.perf_counter() - t0:.1f} s)")
    out = {}
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        with Solver(method=METHOD_TL_RL, precision=prec) as sv:
            sv.add_body(pl.particles, pl.params)
            sv.set_strip_bc(0, pl.left, pl.right, pl.v_bc)
            sv.step(max(warmup, 1))  # includes the one-time neighbour search and correction
            sv.profile(True)
            sv.step(steps)
            prof = sv.profile_read()
            sv.profile(False)
            t0 = time.perf_counter()
            sv.step(steps)
            wall_graph = time.perf_counter() - t0
        out[name] = dict(particle_steps_per_s=n * steps / (prof["step_ms"] / 1e3),
                         particle_steps_per_s_graph=n * steps / wall_graph,
                         ms_per_step=prof["step_ms"] / steps)
        log(f"  {name}: {out[name]['particle_steps_per_s']:.4e} particle-steps/s "
            f"({out[name]['ms_per_step']:.3f} ms/step), graph {out[name]['particle_steps_per_s_graph']:.4e}")
    if accuracy_steps > 0:
        # mixed vs FP64 over the configuration's horizon, from the same state
        states = {}
        for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
            with Solver(method=METHOD_TL_RL, precision=prec) as sv:
                sv.add_body(pl.particles, pl.params)
                sv.set_strip_bc(0, pl.left, pl.right, pl.v_bc)
                t0 = time.perf_counter()
                sv.step(accuracy_steps)
                states[name] = sv.state(0)
                log(f"  accuracy run {name}: {accuracy_steps} steps in {time.perf_counter() - t0:.2f} s")
        out["accuracy"] = dict(steps=accuracy_steps, **errors(states["mixed"], states["fp64"]))
        log(f"  mixed vs fp64 after {accuracy_steps} steps: {out['accuracy']}")
    return n, out


def bench_godunov(steps, warmup, log, cpu_steps=1, nbox=1213):
    """Godunov (Parshikov) UL SPH: the reference's rings_godunov scene
    (src/benchmarks/rings.cpp:101-161: h = 1.7 dx, rho0 = 1, +-180 m/s,
    dt = 1e-7 s) at dx = 6.6e-5 m (nbox = 1213, 1,009,204 particles), FP64.
    The pair kernel is timed with HIP events on the context's stream; the
    CPU oracle (one host thread, the reference's sort order) runs cpu_steps
    steps of the same scene, and the GPU state after as many steps is
    compared with it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from mfree_iwf_amd import METHOD_UL_GODUNOV, Solver, scenes
    t0 = time.perf_counter()
    sc = scenes.reference_rings_godunov(nbox)
    n = sc.particles.n
    log(f"godunov: {n} particles (setup {time.perf_counter() - t0:.1f} s)")
    out = {}
    with Solver(method=METHOD_UL_GODUNOV) as sv:
        sv.add_body(sc.particles, sc.params)
        if cpu_steps > 0:
            sv.step(cpu_steps)
            g = sv.state(0)
        sv.step(max(warmup, 1))
        sv.profile(True)
        sv.step(steps)
        prof = sv.profile_read()
        sv.profile(False)
        t0 = time.perf_counter()
        sv.step(steps)
        wall_graph = time.perf_counter() - t0
    pair_ms = prof["pair_ms"] / max(prof["pair_launches"], 1)
    achieved = ALG_BYTES_PER_PARTICLE_STAGE * n / (pair_ms * 1e-3) / 1e9
    out["fp64"] = dict(particle_steps_per_s=n * steps / (prof["step_ms"] / 1e3),
                       particle_steps_per_s_graph=n * steps / wall_graph, ms_per_step=prof["step_ms"] / steps,
                       pair_kernel_ms=pair_ms,
                       roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                                     frac=achieved / HBM_PEAK_GBS, kernel="k_ul_pairs64<MFREE_METHOD_UL_GODUNOV>",
                                     algorithmic_bytes_per_launch=ALG_BYTES_PER_PARTICLE_STAGE * n))
    log(f"  fp64: {out['fp64']['particle_steps_per_s']:.4e} particle-steps/s ({out['fp64']['ms_per_step']:.3f} "
        f"ms/step), pair kernel {pair_ms:.3f} ms")
    if cpu_steps > 0:
        t0 = time.perf_counter()
        rc, o = O.ul_step(sc.particles, sc.params, cpu_steps, order=O.ORDER_REFERENCE, threads=1,
                          method=O.METHOD_GODUNOV)
        dt = time.perf_counter() - t0
        if rc != O.OK:
            raise RuntimeError(f"oracle failed with status {rc}")
        model, ncpu = cpu_model()
        out["cpu_baseline"] = dict(value=n * cpu_steps / dt, unit="particle-steps/s", cores=1, kind="port",
                                   sample=f"{n} particles (the full scene), {cpu_steps} RK4 step(s); SoA oracle "
                                          f"(g++ -O2) in the reference's sort order on 1 thread",
                                   cpu_model=model, host_logical_cpus=ncpu, seconds=dt)
        out["gpu_vs_oracle"] = dict(steps=cpu_steps, **errors(g, o))
        log(f"  cpu baseline: {cpu_steps} step(s) in {dt:.1f} s; GPU vs oracle: {out['gpu_vs_oracle']}")
    return n, out


def bench_contact(steps, warmup, log, nbox=1213):
    """The two-ring contact driver (rings_monaghan_contact,
    src/benchmarks/rings.cpp:522-620) as two bodies of 504,602 particles
    each (nbox = 1213, dx = 6.6e-5 m), mixed precision, stepped with the
    reference main loop's cadence: one RK4 step of both bodies, the
    outline update, the penalty contact of both ordered pairs
    (src/mfree_iwf.cpp:84-97).  The driver's dt (1e-6 s, sized for dx =
    1e-3 m) is replaced by the CFL-0.3 step of this resolution, and the
    rings start overlapping by one dx, closing at 360 m/s: the timed steps
    run in contact.  Reports the loop's rate and the time of the outline update
    and contact kernels per step (the loop's time less that of plain
    steps)."""
    from mfree_iwf_amd import PRECISION_MIXED, Solver, convex_hull, scenes
    t0 = time.perf_counter()
    dx = 0.08 / (nbox - 1)
    dt = scenes.cfl_dt(0.3, 1.7 * dx, 1e7, 0.4, 1.0, 180.0)
    bodies = scenes.reference_rings_contact(nbox, dt=dt)
    gap = bodies[1].particles.x.min() - bodies[0].particles.x.max()
    for k, b in enumerate(bodies):
        b.particles.x += (0.5 * (gap + b.dx)) * (1 if k == 0 else -1)
    n = sum(b.particles.n for b in bodies)
    hulls = [convex_hull(b.particles.x, b.particles.y) for b in bodies]
    log(f"contact: 2 bodies, {n} particles, outlines {len(hulls[0])} + {len(hulls[1])} points "
        f"(setup {time.perf_counter() - t0:.1f} s)")
    with Solver(precision=PRECISION_MIXED) as sv:
        for k, b in enumerate(bodies):
            sv.add_body(b.particles, b.params)
            sv.set_boundary(k, hulls[k])

        def loop(k):
            for _ in range(k):
                sv.step(1)
                sv.update_boundaries()
                sv.contact_all()
        loop(max(warmup, 1))
        sv.state(0)  # synchronise
        t0 = time.perf_counter()
        loop(steps)
        sv.state(0)
        wall = time.perf_counter() - t0
        t0 = time.perf_counter()
        sv.step(steps)
        sv.state(0)
        wall_step = time.perf_counter() - t0
        fx, _ = sv.contact_forces(0)
    out = dict(particles=n, particle_steps_per_s=n * steps / wall, ms_per_step=wall / steps * 1e3,
               ms_per_step_without_contact=wall_step / steps * 1e3,
               contact_ms_per_step=(wall - wall_step) / steps * 1e3,
               outline_points=[int(len(h)) for h in hulls], particles_in_contact=int(np.count_nonzero(fx)))
    log(f"  {out['particle_steps_per_s']:.4e} particle-steps/s ({out['ms_per_step']:.3f} ms/step, contact "
        f"{out['contact_ms_per_step']:.3f} ms), {out['particles_in_contact']} particles of body 0 in contact")
    return n, out


def bench_tl_variants(steps, warmup, log):
    """The co-rotated and potential TL action lists (tensile_corot,
    tensile_potential) on the tension plate with a hole (0.1 x 0.04 m,
    r = 0.01 m, dx = 8e-5 m, 577,695 particles), FP64.  The potential list's
    forces are the gradient of its potential divided by a particle volume,
    which at this resolution (V = 6.4e-9 m^2) needs a time step far below
    the CFL step: it runs at 1e-6 of it (the rate does not depend on dt)."""
    from mfree_iwf_amd import METHOD_TL_COROTATIONAL, METHOD_TL_POTENTIAL, Solver, scenes
    out = {}
    for name, method in (("corot", METHOD_TL_COROTATIONAL), ("potential", METHOD_TL_POTENTIAL)):
        pl = scenes.plate_config3()
        if method == METHOD_TL_POTENTIAL:
            pl.params.dt = pl.params.dt * 1e-6
        n = pl.particles.n
        with Solver(method=method) as sv:
            sv.add_body(pl.particles, pl.params)
            sv.set_strip_bc(0, pl.left, pl.right, pl.v_bc)
            sv.step(max(warmup, 1))
            sv.profile(True)
            sv.step(steps)
            prof = sv.profile_read()
            sv.profile(False)
        out[name] = dict(particles=n, particle_steps_per_s=n * steps / (prof["step_ms"] / 1e3),
                         ms_per_step=prof["step_ms"] / steps)
        log(f"  {name}: {out[name]['particle_steps_per_s']:.4e} particle-steps/s ({out[name]['ms_per_step']:.3f} ms/step)")
    return out


def traffic_from_profiles(kernel, particles):
    """The committed rocprof traffic figure of this kernel at this size, or None."""
    p = os.path.join(ROOT, "profiles", "pair_kernel_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    for e in d.get("entries", [d]):
        if e.get("kernel") == kernel and e.get("particles") == particles:
            return e
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run the plain copy kernel over MIB mebibytes (HBM counter calibration)")
    ap.add_argument("--modes", default="plain_fp64,fp64,mixed,surrogate")
    ap.add_argument("--plain-steps", type=int, default=10,
                    help="timed steps of the reference-structured FP64 baseline")
    ap.add_argument("--sur-scenes", type=int, default=256)
    ap.add_argument("--sur-captures", type=int, default=80)
    ap.add_argument("--sur-gap", type=int, default=0)
    ap.add_argument("--sur-iters", type=int, default=8000)
    ap.add_argument("--sur-order", type=int, default=4, choices=[3, 4],
                    help="base step of a learned step: RK4 (4) or Kutta's third-order method (3) over 4 dt")
    ap.add_argument("--sur-collect", default="fp64", choices=["fp64", "mixed"],
                    help="precision of the batched solver that generates the training rollouts")
    ap.add_argument("--sur-lr", type=float, default=1e-3)
    ap.add_argument("--sur-batch", type=int, default=1 << 20)
    ap.add_argument("--sur-group", type=int, default=16)
    ap.add_argument("--sur-threshold", type=float, default=1e-4)
    ap.add_argument("--sur-steps", type=int, default=0,
                    help="rollout length (default: the accuracy horizon, --accuracy-steps)")
    ap.add_argument("--workload", default="rings1m",
                    choices=["rings1m", "batch256", "plate578k", "godunov1m", "contact1m", "tl_variants578k"])
    ap.add_argument("--validate-fp64-steps", type=int, default=0,
                    help="only run the FP64-vs-oracle gate over this many steps and print its JSON")
    ap.add_argument("--no-extra-workloads", action="store_true",
                    help="skip the compact batch256, plate578k and godunov1m results of the default line")
    ap.add_argument("--oracle-state", default="",
                    help="npz of the oracle's state of the workload after --accuracy-steps steps: errors of every "
                         "mode against it are added")
    args = ap.parse_args()
    if args.oracle_state:
        global ORACLE_STATE
        ORACLE_STATE = load_oracle_state(args.oracle_state)

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, scenes
    if args.workload == "batch256":
        # bodies are isolated: the fastest impacts (up to Mach 2) exceed the
        # reference's 200-neighbour limit after about 100 steps and freeze
        ntot, out = bench_batch(min(args.steps, 20), min(args.warmup, 3), 256, log,
                                accuracy_steps=500 if args.accuracy_steps == 1000 else args.accuracy_steps)
        acc = out.pop("accuracy", None)
        head = out["mixed"]
        print(json.dumps({
            "metric": "particle-steps/s, 256 uneven ring-collision bodies batched on one GPU",
            "value": head["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": min(args.steps, 20), "warmup": min(args.warmup, 3), "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"256 ring scenes from seed 0, {ntot} particles, 2k-60k per scene, "
                                   "mixed precision", "particles": ntot, "mode": "mixed"},
            "accuracy": acc, "modes": out}), flush=True)
        return
    if args.workload == "plate578k":
        n, out = bench_plate(args.steps, args.warmup, log,
                             accuracy_steps=2000 if args.accuracy_steps == 1000 else args.accuracy_steps)
        acc = out.pop("accuracy", None)
        head = out["mixed"]
        print(json.dumps({
            "metric": "particle-steps/s, Total-Lagrangian corrected SPH tension plate",
            "value": head["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"tension plate 0.1 x 0.04 m, hole r = 0.01 m, dx = 8e-5 m, {n} particles, "
                                   "strips 3 dx at +-1 m/s, mixed precision", "particles": n, "mode": "mixed"},
            "accuracy": acc, "modes": out}), flush=True)
        return
    if args.workload in ("contact1m", "tl_variants578k"):
        if args.workload == "contact1m":
            n, out = bench_contact(min(args.steps, 20), min(args.warmup, 3), log)
            head, metric, cfg = out, "particle-steps/s, two-ring contact (penalty contact between bodies)", dict(
                workload=f"rings_monaghan_contact scene, 2 bodies, {n} particles, mixed precision, main-loop "
                         "cadence (step, outline update, contact)", particles=n, mode="mixed")
        else:
            out = bench_tl_variants(min(args.steps, 20), min(args.warmup, 3), log)
            head, metric, cfg = out["corot"], "particle-steps/s, co-rotated TL tension plate", dict(
                workload="tension plate 0.1 x 0.04 m, hole r = 0.01 m, dx = 8e-5 m, co-rotated and potential "
                         "TL lists, FP64", particles=out["corot"]["particles"], mode="fp64")
        print(json.dumps({"metric": metric, "value": head["particle_steps_per_s"], "unit": "particle-steps/s",
                          "n_gpus": 1, "steps": min(args.steps, 20), "warmup": min(args.warmup, 3),
                          "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32+f64" if args.workload == "contact1m" else "f64",
                          "data": "synthetic", "config": cfg, "modes": out}), flush=True)
        return
    if args.workload == "godunov1m":
        n, out = bench_godunov(min(args.steps, 20), min(args.warmup, 3), log,
                               cpu_steps=0 if args.no_cpu_baseline else 1)
        head = out["fp64"]
        print(json.dumps({
            "metric": "particle-steps/s, Godunov (Parshikov) UL SPH ring collision",
            "value": head["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": min(args.steps, 20), "warmup": min(args.warmup, 3), "ms_per_step": head["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"rings_godunov scene, dx = 6.6e-5 m, {n} particles, h = 1.7 dx, FP64",
                       "particles": n, "mode": "fp64"},
            "roofline": head["roofline"], "cpu_baseline": out.get("cpu_baseline"),
            "gpu_vs_oracle": out.get("gpu_vs_oracle"), "modes": out}), flush=True)
        return
    t0 = time.perf_counter()
    scene = scenes.rings_config2()
    n = scene.particles.n
    log(f"workload: {n} particles, dt = {scene.params.dt:.4e} s (setup {time.perf_counter() - t0:.1f} s)")

    if args.validate_fp64_steps > 0:
        print(json.dumps({"validation": "fp64 GPU vs oracle", **validate_fp64(scene, args.validate_fp64_steps, log)}),
              flush=True)
        return

    calib = None
    if args.calibrate_copy > 0:
        import ctypes as C
        from mfree_iwf_amd import _capi
        g = C.c_double(0)
        _capi.check(_capi.lib().mfree_calibrate_copy(0, args.calibrate_copy << 20, 5, C.byref(g)),
                    what="mfree_calibrate_copy")
        calib = dict(bytes_per_launch_read=args.calibrate_copy << 20, bytes_per_launch_write=args.calibrate_copy << 20,
                     gbps=g.value)
        log(f"copy calibration: {args.calibrate_copy} MiB, {g.value:.1f} GB/s")

    modes = {}
    wanted = args.modes.split(",")
    from mfree_iwf_amd import _capi
    for name, prec, pipe, nst in (("plain_fp64", PRECISION_FP64, _capi.PIPELINE_REFERENCE, args.plain_steps),
                                  ("fp64", PRECISION_FP64, _capi.PIPELINE_FUSED, args.steps),
                                  ("mixed", PRECISION_MIXED, _capi.PIPELINE_FUSED, args.steps)):
        if name not in wanted:
            continue
        log(f"mode {name}")
        r = run_mode(scene, prec, nst, min(args.warmup, nst), log, pipeline=pipe)
        r.pop("state")
        modes[name] = r
        log(f"  {r['particle_steps_per_s']:.4e} particle-steps/s ({r['ms_per_step']:.3f} ms/step), "
            f"graph {r['particle_steps_per_s_graph']:.4e}, pair kernel {r['pair_kernel_ms']:.3f} ms")

    acc, ref_state = accuracy(scene, args.accuracy_steps, log) if args.accuracy_steps > 0 else (None, None)
    if acc:
        log(f"  mixed vs fp64 after {acc['steps']} steps: pos {acc['rel_l2_pos']:.3e} "
            f"stress {acc['rel_l2_stress']:.3e}")
    if "surrogate" in wanted:
        log("mode surrogate")
        modes["surrogate"] = bench_surrogate(scene, args, ref_state, log)

    if "plain_fp64" in modes:
        base = modes["plain_fp64"]["particle_steps_per_s"]
        for k, r in modes.items():
            r["speedup_vs_plain_fp64"] = r["particle_steps_per_s"] / base
    head = modes.get("mixed") or modes.get("fp64")
    if head is None:  # a surrogate-only run: no SPH pair-kernel timing
        print(json.dumps({"metric": METRIC, "modes": modes, "accuracy": acc}), flush=True)
        return
    achieved = ALG_BYTES_PER_PARTICLE_STAGE * n / (head["pair_kernel_ms"] * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                traffic=None, kernel="k_ul_list_lds32" if head is modes.get("mixed") else "k_ul_pairs64",
                algorithmic_bytes_per_launch=ALG_BYTES_PER_PARTICLE_STAGE * n)
    tr = traffic_from_profiles(roof["kernel"], n)
    if tr:
        roof["traffic"] = tr.get("hbm_bytes_per_launch")
    # the arithmetic ceiling beside it: the frozen flop count of one stage
    # (128 per active pair, 100 per particle, 20.85 active pairs per particle
    # on this workload) against the vector peak of the pair terms' type
    flops = N_FLOP_PER_PARTICLE_STAGE * n
    af = flops / (head["pair_kernel_ms"] * 1e-3) / 1e12
    vpeak = FP32_VALU_PEAK_TFLOPS if head is modes.get("mixed") else FP64_VALU_PEAK_TFLOPS
    roof_valu = dict(bound="valu", achieved=af, peak=vpeak, unit="TFLOP/s", frac=af / vpeak, kernel=roof["kernel"],
                     algorithmic_flops_per_launch=flops)

    cpu = None
    if rank == 0 and not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, args.cpu_steps, log)

    extra = None
    if not args.no_extra_workloads and "mixed" in modes:
        # the other configurations, compact: 256 uneven bodies (20 timed
        # steps, 500-step horizon) and the TL tension plate (50 timed steps,
        # 2000-step horizon)
        extra = {}
        log("workload batch256")
        ntot, b = bench_batch(20, 3, 256, log, accuracy_steps=500)
        extra["batch256"] = dict(particles=ntot, scenes=256, unit="particle-steps/s", **b)
        log("workload plate578k")
        npl, p = bench_plate(min(args.steps, 50), min(args.warmup, 5), log, accuracy_steps=2000)
        extra["plate578k"] = dict(particles=npl, unit="particle-steps/s", **p)
        log("workload godunov1m")
        ngd, gd = bench_godunov(20, 3, log, cpu_steps=0 if args.no_cpu_baseline else 1)
        extra["godunov1m"] = dict(particles=ngd, unit="particle-steps/s", **gd)
        log("workload contact1m")
        _, ct = bench_contact(20, 3, log)
        extra["contact1m"] = dict(unit="particle-steps/s", **ct)
        log("workload tl_variants578k")
        extra["tl_variants578k"] = dict(unit="particle-steps/s", **bench_tl_variants(20, 3, log))

    line = {
        "metric": METRIC,
        "value": head["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "ring collision, dx=6.6e-5 m, 1,009,204 particles, shuffled (seed 0), "
                               "N(0, 0.01 m/s) jitter, mixed precision", "particles": n,
                   "dt_s": scene.params.dt, "mode": "mixed"},
        "modes": {k: {kk: v for kk, v in r.items()} for k, r in modes.items()},
        "accuracy": acc,
        "roofline": roof,
        "roofline_valu": roof_valu,
        "cpu_baseline": cpu,
        "copy_calibration": calib,
        "workloads": extra,
    }
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
