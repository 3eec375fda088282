"""The reference-shaped C++ surface (include/mfree/: body<T>, actions,
drivers) driving the GPU library: a main written like the reference's,
checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from _contact import OracleBodies
from mfree_iwf_amd import Params, scenes
from mfree_iwf_amd.scenes import ParticleSet

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "reference_main")


def run_driver(*args):
    assert os.path.exists(EXE), "build tests/cpp first (make -C tests/cpp)"
    print(f"[cpp driver] {' '.join(args)}", flush=True)
    out = subprocess.run([EXE, *args], check=True, capture_output=True, text=True, timeout=120).stdout
    lines = out.strip().splitlines()
    n, t = lines[0].split()
    d = np.array([list(map(float, ln.split())) for ln in lines[1:]])
    return int(n), float(t), d


def run_driver_bodies(*args):
    """The driver's output for every body: [(n, time, rows)]."""
    assert os.path.exists(EXE), "build tests/cpp first (make -C tests/cpp)"
    print(f"[cpp driver] {' '.join(args)}", flush=True)
    out = subprocess.run([EXE, *args], check=True, capture_output=True, text=True, timeout=120).stdout
    lines = out.strip().splitlines()
    bodies, k = [], 0
    while k < len(lines):
        n, t = lines[k].split()
        n = int(n)
        bodies.append((n, float(t), np.array([list(map(float, ln.split())) for ln in lines[k + 1:k + 1 + n]])))
        k += 1 + n
    return bodies


def test_rings_contact_driver_main_loop_vs_oracle():
    # rings_monaghan_contact through the reference main loop's cadence in
    # C++ (step + update_boundary per body, contact_apply_body_to_boundary
    # for every ordered pair) against the oracle's loop
    out = run_driver_bodies("rings_contact", "41", "30")
    orc = OracleBodies(scenes.reference_rings_contact(41), "monaghan")
    for _ in range(30):
        orc.step()
    assert len(out) == 2
    for k, (n, t, d) in enumerate(out):
        o = orc.ps[k]
        assert n == o.n
        ep = O.rel_l2(np.concatenate([d[:, 1], d[:, 2]]), np.concatenate([o.x, o.y]))
        ev = O.rel_l2(np.concatenate([d[:, 3], d[:, 4]]), np.concatenate([o.vx, o.vy]))
        ef = O.rel_l2(np.concatenate([d[:, 9], d[:, 10]]), np.concatenate(orc.fc[k]))
        assert ep <= 1e-10 and ev <= 1e-10 and ef <= 1e-8, (k, ep, ev, ef)
        assert np.count_nonzero(d[:, 9]) > 0
        # every body's step advances the global time (src/rk4.h:107)
        assert abs(t - 60e-6) < 1e-15


def test_rings_monaghan_driver_vs_oracle():
    n, t, d = run_driver("rings", "41", "20")
    s = scenes.reference_rings(41)
    assert n == s.particles.n
    rc, o = O.ul_step(s.particles, s.params, 20, order=O.ORDER_STABLE)
    assert rc == O.OK
    ep = O.rel_l2(np.concatenate([d[:, 1], d[:, 2]]), np.concatenate([o.x, o.y]))
    es = O.rel_l2(np.concatenate([d[:, 6], d[:, 7], d[:, 8]]), np.concatenate([o.Sxx, o.Sxy, o.Syy]))
    assert ep <= 1e-10 and es <= 1e-10, (ep, es)
    assert abs(t - 20e-6) < 1e-15


def tensile_reference_scene(nbox):
    dx = 1.0 / (nbox - 1)
    n = nbox * nbox
    ps = ParticleSet.zeros(n)
    ii, jj = np.meshgrid(np.arange(nbox), np.arange(nbox), indexing="ij")
    ps.x[:] = (ii * dx).ravel()
    ps.y[:] = (jj * dx).ravel()
    ps.X[:] = ps.x
    ps.Y[:] = ps.y
    ps.rho[:] = 1.0
    ps.h[:] = dx
    ps.m[:] = dx * dx
    left = np.nonzero(ii.ravel() == 0)[0]
    right = np.nonzero(ii.ravel() == nbox - 1)[0]
    prm = Params(E=1e7, nu=0.4, rho0=1.0, wdeltap=0.0, corr_exp=0.0, mghn_eps=0.0, alpha=1.0, beta=1.0, eta=0.1,
                 xsph_eps=0.0, dt=1e-6, time=0.0)
    return ps, prm, left, right, dx


def test_tensile_tl_driver_vs_oracle():
    n, t, d = run_driver("tensile", "21", "20")
    ps, prm, left, right, dx = tensile_reference_scene(21)
    assert n == ps.n
    tl = O.TL(ps, dx, search=1, contrib_only=1)
    assert tl.rc == O.OK
    tl.set_bc(left, right, 10.0)
    rc, o, _ = tl.step(ps, prm, 20)
    assert rc == O.OK
    ep = O.rel_l2(np.concatenate([d[:, 1], d[:, 2]]), np.concatenate([o.x, o.y]))
    es = O.rel_l2(np.concatenate([d[:, 6], d[:, 7], d[:, 8]]), np.concatenate([o.Sxx, o.Sxy, o.Syy]))
    assert ep <= 1e-10 and es <= 1e-10, (ep, es)
    assert np.abs(o.Sxx).max() > 0


def test_rings_driver_writes_the_reference_results_files(tmp_path):
    """vtk_writer_write (include/mfree/vtk_writer.h) in the reference's main
    loop cadence: results/out_%06d.vtk after steps 0, 10, 20; the last file
    holds the state the driver prints after its 21 steps."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
    from vtk_format import vtk_text
    assert os.path.exists(EXE), "build tests/cpp first (make -C tests/cpp)"
    (tmp_path / "results").mkdir()
    out = subprocess.run([EXE, "rings", "41", "21", "vtk"], cwd=tmp_path, check=True, capture_output=True, text=True,
                         timeout=120).stdout
    assert sorted(os.listdir(tmp_path / "results")) == ["out_000000.vtk", "out_000010.vtk", "out_000020.vtk"]
    lines = out.strip().splitlines()
    d = np.array([list(map(float, ln.split())) for ln in lines[1:]])
    s = scenes.reference_rings(41)
    st = ParticleSet.zeros(s.particles.n)
    st.h[:] = s.particles.h  # h is constant in UL
    for k, f in enumerate(("x", "y", "vx", "vy", "rho", "Sxx", "Sxy", "Syy")):
        getattr(st, f)[:] = d[:, 1 + k]
    assert (tmp_path / "results" / "out_000020.vtk").read_text() == vtk_text([st])
