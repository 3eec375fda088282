"""Pythonic handle on one mfree context (a set of bodies stepped together).

Thin wrapper over the C ABI (include/mfree_iwf_amd.h); all compute runs in
the gfx950 library.
"""
import ctypes as C

import numpy as np

from . import _capi
from .scenes import ParticleSet


class Solver:
    def __init__(self, method=_capi.METHOD_UL_MONAGHAN, precision=_capi.PRECISION_FP64, use_graph=True,
                 deterministic=True, max_nbh=200, device=0, tl_h_search=0.0, tile_cells=0,
                 pipeline=_capi.PIPELINE_FUSED, xcd_chunk=0, cell_scale=0.0, neighbor_list=0, list_lanes=0,
                 isolate_bodies=False, list_staging=0):
        L = _capi.lib()
        opt = _capi.Options()
        L.mfree_default_options(C.byref(opt))
        opt.method = method
        opt.precision = precision
        opt.use_graph = 1 if use_graph else 0
        opt.deterministic = 1 if deterministic else 0
        opt.max_nbh = max_nbh
        opt.device = device
        opt.tl_h_search = tl_h_search
        opt.tile_cells = tile_cells
        opt.pipeline = pipeline
        opt.xcd_chunk = xcd_chunk
        opt.cell_scale = cell_scale
        opt.neighbor_list = neighbor_list
        opt.list_lanes = list_lanes
        opt.isolate_bodies = 1 if isolate_bodies else 0
        opt.list_staging = list_staging
        h = C.c_void_p()
        _capi.check(L.mfree_create(C.byref(opt), C.byref(h)), what="mfree_create")
        self._h = h
        self._L = L
        self.sizes = []
        self._params = []

    # -- lifecycle
    def close(self):
        if self._h:
            self._L.mfree_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, what):
        _capi.check(st, self._h, what)

    # -- bodies
    def add_body(self, particles: ParticleSet, params: _capi.Params) -> int:
        cp = particles.to_c()
        bid = C.c_uint32(0)
        self._chk(self._L.mfree_add_body(self._h, C.byref(params), C.byref(cp), particles.n, C.byref(bid)),
                  "mfree_add_body")
        self.sizes.append(particles.n)
        self._params.append(params)
        return bid.value

    def set_strip_bc(self, body, left, right, v_bc):
        left = np.ascontiguousarray(left, dtype=np.uint32)
        right = np.ascontiguousarray(right, dtype=np.uint32)
        self._chk(self._L.mfree_set_strip_bc(self._h, body, left.ctypes.data_as(_capi.u32ptr), len(left),
                                             right.ctypes.data_as(_capi.u32ptr), len(right), v_bc),
                  "mfree_set_strip_bc")

    def tl_precompute(self):
        self._chk(self._L.mfree_tl_precompute(self._h), "mfree_tl_precompute")

    @property
    def n_total(self):
        return int(sum(self.sizes))

    # -- stepping
    def step(self, n=1):
        self._chk(self._L.mfree_step(self._h, n), "mfree_step")

    def state(self, body=0) -> ParticleSet:
        ps = ParticleSet.zeros(self.sizes[body])
        cp = ps.to_c()
        self._chk(self._L.mfree_get_state(self._h, body, C.byref(cp)), "mfree_get_state")
        return ps

    def set_state(self, body, ps: ParticleSet):
        cp = ps.to_c()
        self._chk(self._L.mfree_set_state(self._h, body, C.byref(cp)), "mfree_set_state")

    def time(self, body=0):
        t = C.c_double(0)
        self._chk(self._L.mfree_get_time(self._h, body, C.byref(t)), "mfree_get_time")
        return t.value

    def body_status(self, body=0):
        """0 for a body that steps; the error status that froze it in a
        context created with isolate_bodies=True."""
        st = C.c_int32(0)
        self._chk(self._L.mfree_body_status(self._h, body, C.byref(st)), "mfree_body_status")
        return st.value

    def dt(self, body=0):
        return self._params[body].dt

    def write_vtk(self, path, wait=True):
        """Writes every body's current state as a legacy VTK file (the
        reference's results format).  wait=False returns once the snapshot
        is taken; the file is written while later steps run."""
        import os
        f = self._L.mfree_vtk_write if wait else self._L.mfree_vtk_write_async
        self._chk(f(self._h, os.fsencode(path)), "mfree_vtk_write")

    def vtk_wait(self):
        self._chk(self._L.mfree_vtk_wait(self._h), "mfree_vtk_wait")

    def failed_bodies(self):
        return {k: s for k in range(len(self.sizes)) if (s := self.body_status(k)) != 0}

    def energy(self, body=0):
        """FP64 energy of a body: (kinetic, volumetric, deviatoric) in J per
        unit thickness (see mfree_energy)."""
        e = np.zeros(3)
        self._chk(self._L.mfree_energy(self._h, body, e.ctypes.data_as(_capi.dptr)), "mfree_energy")
        return float(e[0]), float(e[1]), float(e[2])

    # -- penalty contact between bodies (src/contact.h)
    def set_boundary(self, body, idx):
        """The body's outline: particle idx in order around it (the drivers
        take the convex hull, :func:`convex_hull`)."""
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self._chk(self._L.mfree_set_boundary(self._h, body, idx.ctypes.data_as(_capi.u32ptr), len(idx)),
                  "mfree_set_boundary")

    def update_boundaries(self):
        self._chk(self._L.mfree_update_boundaries(self._h), "mfree_update_boundaries")

    def boundary(self, body=0):
        """(n, 4) array of the body's outline points: x, y, nx, ny."""
        n = C.c_uint32(0)
        self._chk(self._L.mfree_get_boundary(self._h, body, None, 0, C.byref(n)), "mfree_get_boundary")
        pts = np.zeros((n.value, 4))
        self._chk(self._L.mfree_get_boundary(self._h, body, pts.ctypes.data_as(_capi.dptr), n.value, C.byref(n)),
                  "mfree_get_boundary")
        return pts

    def contact(self, master_body, slave_body, master=None):
        """contact_apply_body_to_boundary(master, slave); master is another
        Solver (on the same device) or this one."""
        m = self if master is None else master
        self._chk(self._L.mfree_contact(m._h, master_body, self._h, slave_body), "mfree_contact")

    def contact_all(self):
        self._chk(self._L.mfree_contact_all(self._h), "mfree_contact_all")

    def contact_forces(self, body=0):
        """(fcx, fcy) of the body's particles, idx order."""
        n = self.sizes[body]
        fx, fy = np.zeros(n), np.zeros(n)
        p = _capi.Particles()
        p.fcx = fx.ctypes.data_as(_capi.dptr)
        p.fcy = fy.ctypes.data_as(_capi.dptr)
        self._chk(self._L.mfree_get_state(self._h, body, C.byref(p)), "mfree_get_state")
        return fx, fy

    # -- parity helpers
    def derivatives(self, with_aux=True):
        n = self.n_total
        d8 = np.zeros((n, 8))
        aux = np.zeros((n, 16)) if with_aux else None
        self._chk(self._L.mfree_eval_derivatives(self._h, d8.ctypes.data_as(_capi.dptr),
                                                 aux.ctypes.data_as(_capi.dptr) if with_aux else None),
                  "mfree_eval_derivatives")
        return d8, aux

    def pairs(self, body=0):
        n = self.sizes[body]
        row = np.zeros(n + 1, dtype=np.uint32)
        tot = C.c_uint64(0)
        st = self._L.mfree_get_pairs(self._h, body, row.ctypes.data_as(_capi.u32ptr), None, 0, C.byref(tot))
        if st not in (_capi.OK, _capi.ERR_CAPACITY):
            self._chk(st, "mfree_get_pairs")
        col = np.zeros(max(tot.value, 1), dtype=np.uint32)
        self._chk(self._L.mfree_get_pairs(self._h, body, row.ctypes.data_as(_capi.u32ptr),
                                          col.ctypes.data_as(_capi.u32ptr), col.size, C.byref(tot)),
                  "mfree_get_pairs")
        return row, col[:tot.value]

    @property
    def tile_cells(self):
        """Cells per LDS tile of the mixed pair pass (0: no tiles)."""
        return int(self._L.mfree_tile_cells(self._h))

    @property
    def cell_scale(self):
        """Cell side of the mixed pair pass in units of 2 h_max."""
        return float(self._L.mfree_cell_scale(self._h))

    # -- profiling
    def profile(self, on=True):
        self._chk(self._L.mfree_profile_enable(self._h, 1 if on else 0), "mfree_profile_enable")

    def profile_read(self):
        pm, pn, sm, sn = C.c_double(), C.c_uint64(), C.c_double(), C.c_uint64()
        self._chk(self._L.mfree_profile_read(self._h, C.byref(pm), C.byref(pn), C.byref(sm), C.byref(sn)),
                  "mfree_profile_read")
        return dict(pair_ms=pm.value, pair_launches=pn.value, step_ms=sm.value, steps=sn.value)


def device_count():
    return int(_capi.lib().mfree_device_count())


def convex_hull(x, y):
    """convex_hull (src/convex_hull.h:62-95) by the library's host code: the
    idx of the hull of the points, counter-clockwise from the lowest-leftmost
    point, collinear points kept."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(2 * len(x) + 1, dtype=np.uint32)
    k = _capi.lib().mfree_convex_hull(x.ctypes.data_as(_capi.dptr), y.ctypes.data_as(_capi.dptr), len(x),
                                      out.ctypes.data_as(_capi.u32ptr))
    return out[:k].copy()
