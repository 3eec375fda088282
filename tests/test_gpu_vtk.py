"""GPU checks of the results writer: the reference's main-loop cadence
(src/mfree_iwf.cpp:82-106, a file after every 10th step) with asynchronous
snapshots, against the oracle's restatement of the writer on the states the
context reports, and against the oracle's FP64 states."""
import io
import os
import sys

import numpy as np
import pytest

import _oracle as O
from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver, scenes, vtk

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
from vtk_format import vtk_text  # noqa: E402
from test_vtk import read_legacy_vtk  # noqa: E402

pytestmark = pytest.mark.gpu


def _scenes():
    return [scenes.rings_config1(), scenes.make_rings(31, centre=0.06, shuffle=True, seed=5, jitter=0.3)]


def _solver(bodies, precision=PRECISION_FP64):
    sv = Solver(precision=precision)
    for b in bodies:
        sv.add_body(b.particles, b.params)
    return sv


def test_main_loop_cadence_and_files(tmp_path):
    bodies = _scenes()
    out = io.StringIO()
    with _solver(bodies) as sv:
        paths = vtk.run(sv, 21, freq=10, outdir=tmp_path, out=out)
    names = [os.path.basename(p) for p in paths]
    assert names == ["out_000000.vtk", "out_000010.vtk", "out_000020.vtk"]
    lines = out.getvalue().splitlines()
    assert [int(s.split(":")[0]) for s in lines] == list(range(21))
    # file of step s = the state after s + 1 steps (the loop steps, then writes)
    with _solver(bodies) as ref:
        done = 0
        for s, p in zip((0, 10, 20), paths):
            ref.step(s + 1 - done)
            done = s + 1
            want = vtk_text([ref.state(k) for k in range(len(bodies))])
            with open(p) as f:
                assert f.read() == want, p


def test_file_values_match_the_oracle_state(tmp_path):
    s = scenes.rings_config1()
    with _solver([s]) as sv:
        sv.step(11)
        p = tmp_path / "out_000010.vtk"
        sv.write_vtk(p)
    rc, o = O.ul_step(s.particles, s.params, 11, order=O.ORDER_STABLE)
    assert rc == O.OK
    pts, _, _, sc, vec = read_legacy_vtk(p.read_text())
    tol = lambda a: 5e-7 + 1e-12 * np.abs(a)  # noqa: E731  (6 decimals; FP64 GPU vs oracle ~1e-15)
    assert np.all(np.abs(pts[:, 0] - o.x) <= tol(o.x)) and np.all(np.abs(pts[:, 1] - o.y) <= tol(o.y))
    for name, f in (("h", o.h), ("density", o.rho), ("Sxx", o.Sxx), ("Sxy", o.Sxy), ("Syy", o.Syy)):
        assert np.all(np.abs(sc[name] - f) <= tol(f)), name
    v = vec["velocity"]
    assert np.all(np.abs(v[:, 0] - o.vx) <= tol(o.vx)) and np.all(np.abs(v[:, 1] - o.vy) <= tol(o.vy))


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_async_snapshot_is_the_state_at_the_call(tmp_path, precision):
    bodies = _scenes()
    with _solver(bodies, precision) as sv:
        sv.step(7)
        want = vtk_text([sv.state(k) for k in range(len(bodies))])
        p = tmp_path / "snap.vtk"
        sv.write_vtk(p, wait=False)
        sv.step(40)  # runs while the file is copied and formatted
        q = tmp_path / "snap2.vtk"
        sv.write_vtk(q, wait=False)  # waits for the first file, then snapshots again
        after = vtk_text([sv.state(k) for k in range(len(bodies))])
        sv.vtk_wait()
    assert p.read_text() == want
    assert q.read_text() == after
    assert want != after
