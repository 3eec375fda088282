"""The library's results writer (legacy VTK, src/vtk_writer.h:27-127) from host
arrays: byte for byte against the oracle's restatement of the writer, and
the oracle's text against the legacy VTK grammar with an independent
reader.  No device needed."""
import os
import sys

import numpy as np
import pytest

from mfree_iwf_amd import ParticleSet, scenes, vtk

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
from vtk_format import vtk_text  # noqa: E402


def _bodies():
    a = scenes.make_rings(21, shuffle=True, seed=1, jitter=0.3).particles
    b = scenes.make_rings(17, ri=0.0, centre=0.1, seed=2, jitter=0.5).particles
    rng = np.random.default_rng(3)
    for ps in (a, b):
        ps.vx += rng.normal(0.0, 3.0, ps.n)
        ps.Sxx[:] = rng.normal(0.0, 2e5, ps.n)
        ps.Sxy[:] = rng.normal(0.0, 2e5, ps.n)
        ps.Syy[:] = rng.normal(0.0, 2e5, ps.n)
        ps.rho += rng.normal(0.0, 5.0, ps.n)
    return [a, b]


def read_legacy_vtk(text):
    """A reader of the legacy VTK unstructured-grid subset (vertex cells,
    point scalars and one point vector); returns (points, cells, types,
    scalars, vectors) and checks the counts on every keyword line."""
    lines = text.split("\n")
    assert lines[0] == "# vtk DataFile Version 2.0"
    assert lines[2] == "ASCII" and lines[4] == "DATASET UNSTRUCTURED_GRID"
    k = 5
    kw, n, typ = lines[k].split()
    assert kw == "POINTS" and typ == "float"
    n = int(n)
    pts = np.array([[float(t) for t in lines[k + 1 + i].split()] for i in range(n)]).reshape(n, 3)
    k += 1 + n
    assert lines[k] == ""
    kw, nc, size = lines[k + 1].split()
    assert kw == "CELLS" and int(nc) == n and int(size) == 2 * n
    cells = np.array([[int(t) for t in lines[k + 2 + i].split()] for i in range(n)]).reshape(n, 2)
    k += 2 + n
    assert lines[k] == ""
    kw, nt = lines[k + 1].split()
    assert kw == "CELL_TYPES" and int(nt) == n
    types = np.array([int(lines[k + 2 + i]) for i in range(n)])
    k += 2 + n
    assert lines[k] == ""
    assert lines[k + 1] == f"POINT_DATA {n}"
    k += 2
    scalars = {}
    while lines[k].startswith("SCALARS"):
        _, name, typ, ncomp = lines[k].split()
        assert typ == "float" and ncomp == "1" and lines[k + 1] == "LOOKUP_TABLE default"
        scalars[name] = np.array([float(lines[k + 2 + i]) for i in range(n)])
        k += 2 + n
        assert lines[k] == ""
        k += 1
    _, vname, typ = lines[k].split()
    assert typ == "float"
    vec = np.array([[float(t) for t in lines[k + 1 + i].split()] for i in range(n)]).reshape(n, 3)
    k += 1 + n
    assert lines[k:] == ["", ""]  # the closing empty line, then the end of the text
    return pts, cells, types, scalars, {vname: vec}


def test_oracle_text_is_a_legacy_vtk_grid_of_the_state():
    bodies = _bodies()
    pts, cells, types, sc, vec = read_legacy_vtk(vtk_text(bodies))
    cat = {f: np.concatenate([getattr(b, f) for b in bodies]) for f in ("x", "y", "h", "rho", "Sxx", "Sxy", "Syy",
                                                                          "vx", "vy")}
    n = len(cat["x"])
    # %f keeps 6 decimals: every value within 5e-7 (plus the value's own ulp)
    tol = lambda a: 5e-7 + 1e-15 * np.abs(a)  # noqa: E731
    assert np.all(np.abs(pts[:, 0] - cat["x"]) <= tol(cat["x"]))
    assert np.all(np.abs(pts[:, 1] - cat["y"]) <= tol(cat["y"]))
    assert np.all(pts[:, 2] == 0.0)
    assert np.array_equal(cells, np.stack([np.ones(n, int), np.arange(n)], axis=1))
    assert np.all(types == 1)  # VTK_VERTEX
    assert list(sc) == ["h", "density", "Sxx", "Sxy", "Syy"]
    for name, f in (("h", "h"), ("density", "rho"), ("Sxx", "Sxx"), ("Sxy", "Sxy"), ("Syy", "Syy")):
        assert np.all(np.abs(sc[name] - cat[f]) <= tol(cat[f])), name
    v = vec["velocity"]
    assert np.all(np.abs(v[:, 0] - cat["vx"]) <= tol(cat["vx"])) and np.all(v[:, 2] == 0.0)
    assert np.all(np.abs(v[:, 1] - cat["vy"]) <= tol(cat["vy"]))


def test_library_writer_equals_oracle_text(tmp_path):
    bodies = _bodies()
    p = tmp_path / "out_000000.vtk"
    vtk.write_host(p, bodies)
    assert p.read_text() == vtk_text(bodies)
    one = tmp_path / "one.vtk"
    vtk.write_host(one, bodies[:1])
    assert one.read_text() == vtk_text(bodies[:1])


def test_library_writer_rounding_cases(tmp_path):
    # ties at the 7th decimal (0.0078125 = 1/128), negative zero, values that
    # print as -0.000000, large and tiny magnitudes
    vals = np.array([0.0078125, -0.0078125, -0.0, 0.0, -1e-7, 4e-7, 5e-7, 6e-7, 1.25e7, -3.5e12, 2.5e-6,
                     0.1 + 0.2, 1.0 / 3.0, 123456.7890125])
    ps = ParticleSet.zeros(len(vals))
    for f in ("x", "y", "h", "rho", "Sxx", "Sxy", "Syy", "vx", "vy"):
        getattr(ps, f)[:] = np.roll(vals, len(f))
    p = tmp_path / "r.vtk"
    vtk.write_host(p, [ps])
    assert p.read_text() == vtk_text([ps])


def test_library_writer_many_particles_in_pieces(tmp_path):
    # more particles than one formatting piece (65,536): the pieces are joined in order
    n = 150_001
    ps = ParticleSet.zeros(n)
    rng = np.random.default_rng(9)
    for f in ("x", "y", "h", "rho", "Sxx", "Sxy", "Syy", "vx", "vy"):
        getattr(ps, f)[:] = rng.normal(0.0, 10.0, n)
    p = tmp_path / "big.vtk"
    vtk.write_host(p, [ps])
    assert p.read_text() == vtk_text([ps])


def test_library_writer_empty_and_errors(tmp_path):
    empty = ParticleSet.zeros(0)
    p = tmp_path / "e.vtk"
    vtk.write_host(p, [empty])
    assert p.read_text() == vtk_text([empty])
    from mfree_iwf_amd import MfreeError
    with pytest.raises(MfreeError):
        vtk.write_host(tmp_path / "no_such_dir" / "x.vtk", [empty])
