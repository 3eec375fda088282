"""Results files in the reference's legacy VTK format (src/vtk_writer.h:27-127)
and the reference's main loop cadence (src/mfree_iwf.cpp:82-106): step every
body, write results/out_%06d.vtk after every freq-th step, print the step and
the time.  The files are formatted by the library (mfree_vtk_write*)."""
import ctypes as C
import os
import sys

from . import _capi


def write_host(path, particle_sets):
    """One file from host ParticleSets (bodies in order), no device needed."""
    sets = list(particle_sets)
    views = (_capi.Particles * max(len(sets), 1))()
    for k, ps in enumerate(sets):
        views[k] = ps.to_c()
    counts = (C.c_uint32 * max(len(sets), 1))(*[ps.n for ps in sets])
    _capi.check(_capi.lib().mfree_vtk_write_host(os.fsencode(path), len(sets), C.cast(views, C.c_void_p), counts),
                what="mfree_vtk_write_host")


def run(solver, nsteps, freq=10, outdir="results", out=sys.stdout, body=0, contact=False):
    """The reference's main loop for nsteps steps: after step s (counted from
    0) the state is written to outdir/out_%06d.vtk when s % freq == 0, and
    "s: time" is printed.  With contact, every step is followed by the
    outline update and the penalty contact of every ordered pair of bodies
    (src/mfree_iwf.cpp:84-97; the outlines must have been set).  Files are
    written asynchronously while the next steps run; returns the paths
    written."""
    os.makedirs(outdir, exist_ok=True)
    written = []
    t = solver.time(body)
    dt = solver.dt(body)
    s = 0
    while s < nsteps:
        # steps up to the next write (or the end) in one call
        nxt = s if s % freq == 0 else min(nsteps - 1, (s // freq + 1) * freq)
        if contact:
            for _ in range(nxt - s + 1):
                solver.step(1)
                solver.update_boundaries()
                solver.contact_all()
        else:
            solver.step(nxt - s + 1)
        for k in range(s, nxt + 1):
            t = t + dt
            if out is not None:
                print(f"{k}: {t:f}", file=out)
        if nxt % freq == 0:
            p = os.path.join(outdir, f"out_{nxt:06d}.vtk")
            solver.write_vtk(p, wait=False)
            written.append(p)
        s = nxt + 1
    solver.vtk_wait()
    return written
