"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the compiled reference pieces (oracle/_ref/libmfree_ref.so).

Test infrastructure: the checker, never the thing under test.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmfree_ref.so")

ORDER_REFERENCE = 0
ORDER_STABLE = 1
OK = 0
ERR_NBH_OVERFLOW = -3
ERR_NO_NEIGHBOR = -4
ERR_CELLS = -6

dptr = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)


class Consts(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("E", "nu", "rho0", "wdeltap", "corr_exp", "mghn_eps", "alpha",
                                          "beta", "eta", "xsph_eps")]


class Soa(C.Structure):
    _fields_ = [(n, dptr) for n in ("x", "y", "X", "Y", "m", "rho", "h", "vx", "vy", "Sxx", "Sxy", "Syy",
                                    "fcx", "fcy")]


class Geom(C.Structure):
    _fields_ = [("bbmin_x", C.c_double), ("bbmin_y", C.c_double), ("bbmax_x", C.c_double),
                ("bbmax_y", C.c_double), ("cell", C.c_double), ("nx", C.c_uint32), ("ny", C.c_uint32),
                ("num_cell", C.c_uint32)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_cubic_spline.argtypes = [C.c_double] * 5 + [dptr]
        L.orc_cubic_spline.restype = None
        L.orc_grid_geometry.argtypes = [C.c_uint32, dptr, dptr, dptr, C.c_double, C.POINTER(Geom)]
        L.orc_verlet_pairs.argtypes = [C.c_uint32, dptr, dptr, dptr, C.c_int, C.c_uint32, u32p, u32p,
                                       C.c_uint64, C.POINTER(Geom)]
        L.orc_find_neighbors.argtypes = [C.c_uint32, dptr, dptr, C.c_double, C.c_uint32, u32p, u32p, C.c_uint64]
        L.orc_ul_derivatives.argtypes = [C.POINTER(Consts), C.c_uint32, C.POINTER(Soa), C.c_int, C.c_uint32,
                                         dptr, dptr]
        L.orc_ul_step.argtypes = [C.POINTER(Consts), C.c_uint32, C.POINTER(Soa), C.c_double, C.c_uint32,
                                  C.c_int, C.c_uint32, C.c_int]
        L.orc_ul_derivatives_m.argtypes = [C.POINTER(Consts), C.c_int, C.c_uint32, C.POINTER(Soa), C.c_int,
                                           C.c_uint32, dptr, dptr]
        L.orc_ul_step_m.argtypes = [C.POINTER(Consts), C.c_int, C.c_uint32, C.POINTER(Soa), C.c_double,
                                    C.c_uint32, C.c_int, C.c_uint32, C.c_int]
        L.orc_tl_create.argtypes = [C.c_uint32, dptr, dptr, dptr, dptr, dptr, C.c_double, C.c_int, C.c_int,
                                    C.c_uint32, C.POINTER(C.c_void_p)]
        L.orc_tl_create_m.argtypes = [C.c_int, C.c_uint32, dptr, dptr, dptr, dptr, dptr, C.c_double, C.c_int,
                                      C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]
        L.orc_tl_destroy.argtypes = [C.c_void_p]
        L.orc_tl_num_pairs.argtypes = [C.c_void_p]
        L.orc_tl_num_pairs.restype = C.c_uint64
        L.orc_tl_get.argtypes = [C.c_void_p, u32p, u32p, dptr]
        L.orc_tl_get_moments.argtypes = [C.c_void_p, dptr, dptr]
        L.orc_tl_set_bc.argtypes = [C.c_void_p, u32p, C.c_uint32, u32p, C.c_uint32, C.c_double]
        L.orc_tl_derivatives.argtypes = [C.POINTER(Consts), C.c_void_p, C.c_uint32, C.POINTER(Soa), C.c_double,
                                         dptr, dptr]
        L.orc_tl_step.argtypes = [C.POINTER(Consts), C.c_void_p, C.c_uint32, C.POINTER(Soa), C.c_double,
                                  C.c_uint32, dptr, C.c_int]
        L.orc_convex_hull.argtypes = [dptr, dptr, C.c_uint32, u32p]
        L.orc_convex_hull.restype = C.c_uint32
        L.orc_boundary_update.argtypes = [C.c_uint32, u32p, dptr, dptr, dptr]
        L.orc_contact.argtypes = [C.c_uint32, dptr, C.c_uint32, dptr, dptr, dptr, C.c_double, dptr, dptr]
        L.orc_G.restype = C.c_double
        L.orc_K.restype = C.c_double
        L.orc_G.argtypes = [C.c_double, C.c_double]
        L.orc_K.argtypes = [C.c_double, C.c_double]
        _lib = L
    return _lib


def ref():
    """The compiled reference pieces, or None when they were not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        R.ref_cubic_spline.argtypes = [C.c_double] * 5 + [dptr]
        R.ref_verlet_pairs.argtypes = [C.c_uint32, dptr, dptr, dptr, u32p, u32p, C.c_uint64, dptr]
        R.ref_find_neighbors.argtypes = [C.c_uint32, dptr, dptr, C.c_double, u32p, u32p, C.c_uint64]
        R.ref_convex_hull.argtypes = [dptr, dptr, C.c_uint32, u32p]
        R.ref_convex_hull.restype = C.c_uint32
        R.ref_kd_nearest.argtypes = [dptr, C.c_uint32, dptr, dptr, C.c_uint32, dptr]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(dptr)


def consts_from(params):
    return Consts(E=params.E, nu=params.nu, rho0=params.rho0, wdeltap=params.wdeltap, corr_exp=params.corr_exp,
                  mghn_eps=params.mghn_eps, alpha=params.alpha, beta=params.beta, eta=params.eta,
                  xsph_eps=params.xsph_eps)


def soa_from(ps, fc=None):
    """C view of ps; fc = (fcx, fcy) arrays (idx order) or None.  The
    arrays must stay alive while the view is used."""
    s = Soa()
    for f in ("x", "y", "X", "Y", "m", "rho", "h", "vx", "vy", "Sxx", "Sxy", "Syy"):
        a = getattr(ps, f)
        assert a.dtype == np.float64 and a.flags.c_contiguous
        setattr(s, f, _p(a))
    s.fcx = None
    s.fcy = None
    if fc is not None:
        for a in fc:
            assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (ps.n,)
        s.fcx, s.fcy = _p(fc[0]), _p(fc[1])
    return s


def cubic_spline(xi, yi, xj, yj, h):
    out = (C.c_double * 3)()
    lib().orc_cubic_spline(xi, yi, xj, yj, h, out)
    return out[0], out[1], out[2]


def cubic_spline_ref(xi, yi, xj, yj, h):
    out = (C.c_double * 3)()
    ref().ref_cubic_spline(xi, yi, xj, yj, h, out)
    return out[0], out[1], out[2]


def geometry(ps):
    g = Geom()
    rc = lib().orc_grid_geometry(ps.n, _p(ps.x), _p(ps.y), _p(ps.h), 2.0, C.byref(g))
    return rc, g


def verlet_pairs(x, y, h, order=ORDER_REFERENCE, max_nbh=200, cap=None):
    n = len(x)
    cap = cap if cap is not None else n * 80
    row = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(cap, dtype=np.uint32)
    rc = lib().orc_verlet_pairs(n, _p(x), _p(y), _p(h), order, max_nbh, row.ctypes.data_as(u32p),
                                col.ctypes.data_as(u32p), cap, None)
    return rc, row, col[:row[-1]]


def verlet_pairs_ref(x, y, h, cap=None):
    n = len(x)
    cap = cap if cap is not None else n * 80
    row = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(cap, dtype=np.uint32)
    rc = ref().ref_verlet_pairs(n, _p(x), _p(y), _p(h), row.ctypes.data_as(u32p), col.ctypes.data_as(u32p),
                                cap, None)
    return rc, row, col[:row[-1]]


def find_neighbors(x, y, h, max_nbh=200, cap=None):
    n = len(x)
    cap = cap if cap is not None else n * 80
    row = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(cap, dtype=np.uint32)
    rc = lib().orc_find_neighbors(n, _p(x), _p(y), h, max_nbh, row.ctypes.data_as(u32p),
                                  col.ctypes.data_as(u32p), cap)
    return rc, row, col[:row[-1]]


def find_neighbors_ref(x, y, h, cap=None):
    n = len(x)
    cap = cap if cap is not None else n * 80
    row = np.zeros(n + 1, dtype=np.uint32)
    col = np.zeros(cap, dtype=np.uint32)
    rc = ref().ref_find_neighbors(n, _p(x), _p(y), h, row.ctypes.data_as(u32p), col.ctypes.data_as(u32p), cap)
    return rc, row, col[:row[-1]]


METHOD_MONAGHAN = 0
METHOD_GODUNOV = 1


def ul_derivatives(ps, params, order=ORDER_STABLE, max_nbh=200, method=METHOD_MONAGHAN, fc=None):
    """One UL derivative evaluation (the Gray-Monaghan action list, or the
    Godunov one with method=METHOD_GODUNOV); fc = (fcx, fcy) contact forces."""
    n = ps.n
    d8 = np.zeros((n, 8))
    aux = np.zeros((n, 16))
    c = consts_from(params)
    s = soa_from(ps, fc)
    rc = lib().orc_ul_derivatives_m(C.byref(c), method, n, C.byref(s), order, max_nbh, _p(d8), _p(aux))
    return rc, d8, aux


def ul_step(ps, params, nsteps, order=ORDER_STABLE, threads=1, max_nbh=200, dt=None, method=METHOD_MONAGHAN,
            fc=None):
    """Steps a copy of ps; returns (status, new ParticleSet)."""
    q = ps.copy()
    c = consts_from(params)
    s = soa_from(q, fc)
    rc = lib().orc_ul_step_m(C.byref(c), method, q.n, C.byref(s), params.dt if dt is None else dt, nsteps,
                             order, max_nbh, threads)
    return rc, q


TL_RL = 0
TL_COROTATIONAL = 1
TL_POTENTIAL = 2


class TL:
    """Frozen TL neighbour lists + kernels of the oracle: Randles-Libersky
    (tensile_tl, variant TL_RL), the plain kernel of a co-rotated body
    (TL_COROTATIONAL) or Randles-Libersky with the reverse pairs' kernels of
    a potential body (TL_POTENTIAL)."""

    def __init__(self, ps, h_search, search=1, contrib_only=1, max_nbh=1 << 30, variant=TL_RL):
        self.n = ps.n
        h = C.c_void_p()
        rc = lib().orc_tl_create_m(variant, ps.n, _p(ps.x), _p(ps.y), _p(ps.m), _p(ps.rho), _p(ps.h), h_search,
                                   search, contrib_only, max_nbh, C.byref(h))
        self.rc = rc
        self.h = h if rc == OK else None

    def close(self):
        if self.h:
            lib().orc_tl_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def lists(self):
        npairs = lib().orc_tl_num_pairs(self.h)
        row = np.zeros(self.n + 1, dtype=np.uint32)
        col = np.zeros(npairs, dtype=np.uint32)
        w = np.zeros((npairs, 3))
        lib().orc_tl_get(self.h, row.ctypes.data_as(u32p), col.ctypes.data_as(u32p), _p(w))
        return row, col, w

    def moments(self):
        Cs = np.zeros(self.n)
        B = np.zeros((self.n, 4))
        lib().orc_tl_get_moments(self.h, _p(Cs), _p(B))
        return Cs, B

    def set_bc(self, left, right, v_bc):
        left = np.ascontiguousarray(left, dtype=np.uint32)
        right = np.ascontiguousarray(right, dtype=np.uint32)
        return lib().orc_tl_set_bc(self.h, left.ctypes.data_as(u32p), len(left), right.ctypes.data_as(u32p),
                                   len(right), v_bc)

    def derivatives(self, ps, params, time=0.0, fc=None):
        q = ps.copy()
        d8 = np.zeros((self.n, 8))
        aux = np.zeros((self.n, 16))
        c = consts_from(params)
        s = soa_from(q, fc)
        rc = lib().orc_tl_derivatives(C.byref(c), self.h, self.n, C.byref(s), time, _p(d8), _p(aux))
        return rc, d8, aux, q

    def step(self, ps, params, nsteps, time=0.0, threads=1, dt=None, fc=None):
        q = ps.copy()
        c = consts_from(params)
        s = soa_from(q, fc)
        t = C.c_double(time)
        rc = lib().orc_tl_step(C.byref(c), self.h, self.n, C.byref(s), params.dt if dt is None else dt, nsteps,
                               C.byref(t), threads)
        return rc, q, t.value


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


# ---- penalty contact -------------------------------------------------------

def convex_hull(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(2 * len(x) + 1, dtype=np.uint32)
    k = lib().orc_convex_hull(_p(x), _p(y), len(x), out.ctypes.data_as(u32p))
    return out[:k].copy()


def convex_hull_ref(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.uint32)
    k = ref().ref_convex_hull(_p(x), _p(y), len(x), out.ctypes.data_as(u32p))
    return out[:k].copy()


def boundary_update(idx, x, y):
    """(nb, 4) array of boundary points (x, y, nx, ny)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    pts = np.zeros((len(idx), 4))
    lib().orc_boundary_update(len(idx), idx.ctypes.data_as(u32p), _p(np.ascontiguousarray(x)),
                              _p(np.ascontiguousarray(y)), _p(pts))
    return pts


def contact(pts, x, y, m, dt):
    """Contact forces of slave particles (x, y, m) against master boundary
    points pts; returns (fcx, fcy)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(x)
    fcx, fcy = np.zeros(n), np.zeros(n)
    lib().orc_contact(len(pts), _p(pts), n, _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)),
                      _p(np.ascontiguousarray(m)), dt, _p(fcx), _p(fcy))
    return fcx, fcy


def kd_nearest_ref(pts, qx, qy):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    qx = np.ascontiguousarray(qx, dtype=np.float64)
    qy = np.ascontiguousarray(qy, dtype=np.float64)
    out = np.zeros((len(qx), 4))
    ref().ref_kd_nearest(_p(pts), len(pts), _p(qx), _p(qy), len(qx), _p(out))
    return out

