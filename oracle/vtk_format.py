"""Legacy VTK text of the reference's results writer, restated in Python.

TEST INFRASTRUCTURE ONLY: the checker for the library's VTK writer
(mfree_vtk_write*).  Only the test suite may import it; the library never
uses it.

Restates src/vtk_writer.h:27-127 (vtk_writer_write) line for line: the
header (:38-44), POINTS with z = 0 (:45-51), one vertex cell per particle
(:54-58) and its cell type 1 (:60-64), the point data h, density, Sxx, Sxy,
Syy (:66-115) and the velocity vector (:117-123), every per-section block
followed by an empty line, bodies in order, every value printed as C's
"%f" (Python's "%f" formats doubles identically: 6 decimals, round half to
even on the exact binary value, "-0.000000" for negative zero).

Parity pinning: tests/test_vtk.py checks this text against the legacy VTK
grammar with an independent reader (counts, keywords, cell connectivity,
values within the 5e-7 of a 6-decimal print) and the library's file byte for
byte.
"""


def vtk_text(bodies):
    """bodies: sequence of objects with x, y, h, rho, Sxx, Sxy, Syy, vx, vy
    (1-D float64 arrays of equal length)."""
    n = sum(len(b.x) for b in bodies)
    out = ["# vtk DataFile Version 2.0\n", "mfree iwf\n", "ASCII\n", "\n",
           "DATASET UNSTRUCTURED_GRID\n", "POINTS %d float\n" % n]
    for b in bodies:
        out.extend("%f %f %f\n" % (x, y, 0.0) for x, y in zip(b.x.tolist(), b.y.tolist()))
    out.append("\n")
    out.append("CELLS %d %d\n" % (n, 2 * n))
    out.extend("%d %d\n" % (1, i) for i in range(n))
    out.append("\n")
    out.append("CELL_TYPES %d\n" % n)
    out.extend("%d\n" % 1 for _ in range(n))
    out.append("\n")
    out.append("POINT_DATA %d\n" % n)
    for name, field in (("h", "h"), ("density", "rho"), ("Sxx", "Sxx"), ("Sxy", "Sxy"), ("Syy", "Syy")):
        out.append("SCALARS %s float 1\n" % name)
        out.append("LOOKUP_TABLE default\n")
        for b in bodies:
            out.extend("%f\n" % v for v in getattr(b, field).tolist())
        out.append("\n")
    out.append("VECTORS velocity float\n")
    for b in bodies:
        out.extend("%f %f %f\n" % (vx, vy, 0.0) for vx, vy in zip(b.vx.tolist(), b.vy.tolist()))
    out.append("\n")
    return "".join(out)
