"""Synthetic workloads: parametrised re-creations of the reference drivers.

* ring collision -- ``rings_monaghan`` (src/benchmarks/rings.cpp:20-99): two
  elastic rings on a square lattice, Gray-Monaghan SPH.  The reference's own
  constants are available as :func:`reference_rings`; the benchmark
  configurations use E=1e7, nu=0.4, rho0=1010, +-59 m/s, h=1.3 dx, centres
  +-0.045 m and a CFL-0.3 time step.
* tension plate with a hole -- the Total-Lagrangian tension benchmark
  (src/benchmarks/tensile.cpp:213-289) with a central hole and 3-column
  velocity strips.
* batches of uneven ring scenes drawn from a seed.

Everything is generated by the library's host code (mfree_scene_*), so the
same inputs reach the GPU path, the C++ drivers and the oracle.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import Params, Particles, PARTICLE_FIELDS

STATE_FIELDS = ("x", "y", "vx", "vy", "rho", "Sxx", "Sxy", "Syy")


@dataclass
class ParticleSet:
    """SoA particle arrays (float64) in idx order, fields of src/particle.h:42-60."""
    x: np.ndarray
    y: np.ndarray
    X: np.ndarray
    Y: np.ndarray
    m: np.ndarray
    rho: np.ndarray
    h: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    Sxx: np.ndarray
    Sxy: np.ndarray
    Syy: np.ndarray

    @classmethod
    def zeros(cls, n):
        return cls(**{f: np.zeros(n) for f in cls.__dataclass_fields__})

    @property
    def n(self):
        return int(self.x.shape[0])

    def copy(self):
        return ParticleSet(**{f: getattr(self, f).copy() for f in self.__dataclass_fields__})

    def subset(self, idx):
        return ParticleSet(**{f: np.ascontiguousarray(getattr(self, f)[idx]) for f in self.__dataclass_fields__})

    def to_c(self):
        """C view (the arrays must stay alive while it is used)."""
        p = Particles()
        for f in PARTICLE_FIELDS:
            a = getattr(self, f, None)
            if a is None:
                setattr(p, f, None)
            else:
                assert a.dtype == np.float64 and a.flags.c_contiguous
                setattr(p, f, a.ctypes.data_as(_capi.dptr))
        return p

    def state8(self):
        return np.stack([getattr(self, f) for f in STATE_FIELDS], axis=1)


def cubic_spline(xi, yi, xj, yj, h):
    """The reference's kernel (src/kernel.cpp:20-55), evaluated by the library."""
    out = (C.c_double * 3)()
    _capi.lib().mfree_cubic_spline(xi, yi, xj, yj, h, out)
    return out[0], out[1], out[2]


def elastic_K(E, nu):
    """Bulk modulus with the reference's formula (src/simulation_data.cpp:36-39)."""
    G = E / (2. * (1. + nu))
    return 2.0 * G * (1 + nu) / (3 * (1 - 2 * nu))


def sound_speed(E, nu, rho0):
    return math.sqrt(elastic_K(E, nu) / rho0)


def cfl_dt(cfl, h, E, nu, rho0, vmax):
    """Fixed time step dt = cfl * h / (c0 + max|v|) evaluated at t = 0."""
    return cfl * h / (sound_speed(E, nu, rho0) + vmax)


@dataclass
class RingScene:
    particles: ParticleSet
    params: Params
    dx: float
    nbox: int
    meta: dict = field(default_factory=dict)


def make_rings(nbox, ri=0.03, ro=0.04, centre=0.045, hdx=1.3, rho0=1010.0, v0=59.0, E=1e7, nu=0.4,
               alpha=1.0, beta=1.0, eta=0.1, art_eps=0.3, corr_exp=4.0, xsph_eps=0.5, cfl=0.3, dt=None,
               shuffle=False, seed=0, jitter=0.0):
    spec = _capi.RingsSpec(ri=ri, ro=ro, centre=centre, nbox=nbox, hdx=hdx, rho0=rho0, v0=v0,
                           shuffle=1 if shuffle else 0, seed=seed, jitter=jitter)
    L = _capi.lib()
    n = L.mfree_scene_rings_count(C.byref(spec))
    if n == 0:
        raise ValueError("empty ring scene")
    ps = ParticleSet.zeros(n)
    cp = ps.to_c()
    got = C.c_uint32(0)
    _capi.check(L.mfree_scene_rings(C.byref(spec), C.byref(cp), n, C.byref(got)), what="mfree_scene_rings")
    dx = 2 * ro / (nbox - 1)
    h = hdx * dx
    wdeltap = cubic_spline(0, 0, dx, 0, h)[0]  # src/benchmarks/rings.cpp:74-75
    if dt is None:
        vmax = float(np.sqrt(ps.vx ** 2 + ps.vy ** 2).max())
        dt = cfl_dt(cfl, h, E, nu, rho0, vmax)
    prm = Params(E=E, nu=nu, rho0=rho0, wdeltap=wdeltap, corr_exp=corr_exp, mghn_eps=art_eps,
                 alpha=alpha, beta=beta, eta=eta, xsph_eps=xsph_eps, dt=dt, time=0.0)
    return RingScene(ps, prm, dx, nbox, dict(kind="rings", ri=ri, ro=ro, centre=centre, hdx=hdx, v0=v0,
                                             shuffle=shuffle, seed=seed, jitter=jitter))


def nbox_for_dx(dx, ro=0.04):
    return int(round(2 * ro / dx)) + 1


def rings_config1():
    """Two rings, dx = 1e-3 m (nbox = 81, 4,408 particles)."""
    return make_rings(nbox_for_dx(1e-3))


def rings_config2():
    """Two rings, dx = 6.6e-5 m (nbox = 1213, 1,009,204 particles), shuffled
    from seed 0, N(0, 0.01 m/s) velocity jitter."""
    return make_rings(nbox_for_dx(6.6e-5), shuffle=True, seed=0, jitter=0.01)


def reference_rings(nbox):
    """The reference driver's own constants (src/benchmarks/rings.cpp:22-36,60-80)."""
    return make_rings(nbox, centre=0.04 + 0.001, hdx=1.7, rho0=1.0, v0=180.0, dt=1e-6)


def reference_rings_godunov(nbox, **kw):
    """rings_godunov (src/benchmarks/rings.cpp:101-161): the same lattice and
    material as rings_monaghan, dt = 1e-7 and no corrector constants
    (correction_constants() is all zero); the body is a particle_ul stepped
    with the Godunov action list (method METHOD_UL_GODUNOV)."""
    return make_rings(nbox, centre=0.04 + 0.001, hdx=1.7, rho0=1.0, v0=180.0, dt=1e-7, alpha=0.0, beta=0.0,
                      eta=0.0, art_eps=0.0, corr_exp=0.0, xsph_eps=0.0, **kw)


def reference_rings_contact(nbox, kind="monaghan", dt=None):
    """The two-body contact drivers' scenes (src/benchmarks/rings.cpp:
    rings_monaghan_contact :522-620, rings_godunov_contact :164-243,
    rings_corot_contact :245-335, rings_potential_contact :431-520): the
    rings_monaghan lattice, each ring its own body with particles numbered
    in lattice order, +-180 m/s, E = 1e7, nu = 0.4, rho0 = 1; h = 1.7 dx
    (1.3 dx and rings of radii 0.75 / 1.0 m for the potential driver);
    dt = 1e-6 s for Monaghan, 1e-7 s otherwise; Monaghan's corrections for
    Monaghan bodies, viscosity (1, 1, 0.1) for the TL bodies, none for
    Godunov.  dt overrides the driver's time step (the drivers' steps are
    sized for nbox = 80).  Returns [left, right]; the drivers' outline of
    each is the convex hull of its particles."""
    ri, ro, hdx = (0.75, 1.0, 1.3) if kind == "potential" else (0.03, 0.04, 1.7)
    kw = dict(ri=ri, ro=ro, centre=ro + 0.001, hdx=hdx, rho0=1.0, v0=180.0, E=1e7, nu=0.4,
              dt=dt if dt is not None else (1e-6 if kind == "monaghan" else 1e-7))
    if kind == "godunov":
        kw.update(alpha=0.0, beta=0.0, eta=0.0, art_eps=0.0, corr_exp=0.0, xsph_eps=0.0)
    elif kind in ("corot", "potential"):
        kw.update(art_eps=0.0, corr_exp=0.0, xsph_eps=0.0)
    elif kind != "monaghan":
        raise ValueError(kind)
    s = make_rings(nbox, **kw)
    return [RingScene(s.particles.subset(slice(k, None, 2)), s.params, s.dx, nbox, dict(s.meta, body=k))
            for k in (0, 1)]


@dataclass
class PlateScene:
    particles: ParticleSet
    params: Params
    left: np.ndarray
    right: np.ndarray
    v_bc: float
    dx: float


def make_plate(dx=8e-5, L=0.1, W=0.04, r_hole=0.01, hdx=1.0, rho0=7850.0, E=2e9, nu=0.3, strip=3,
               v_bc=1.0, alpha=1.0, beta=1.0, eta=0.1, cfl=0.3, dt=None):
    spec = _capi.PlateSpec(L=L, W=W, r_hole=r_hole, dx=dx, hdx=hdx, rho0=rho0, strip=strip)
    Lb = _capi.lib()
    nl, nr = C.c_uint32(0), C.c_uint32(0)
    n = Lb.mfree_scene_plate_count(C.byref(spec), C.byref(nl), C.byref(nr))
    if n == 0:
        raise ValueError("empty plate")
    ps = ParticleSet.zeros(n)
    cp = ps.to_c()
    left = np.zeros(max(nl.value, 1), dtype=np.uint32)
    right = np.zeros(max(nr.value, 1), dtype=np.uint32)
    got = C.c_uint32(0)
    _capi.check(Lb.mfree_scene_plate(C.byref(spec), C.byref(cp), n, left.ctypes.data_as(_capi.u32ptr),
                                     right.ctypes.data_as(_capi.u32ptr), C.byref(got)), what="mfree_scene_plate")
    h = hdx * dx
    if dt is None:
        dt = cfl_dt(cfl, h, E, nu, rho0, v_bc)
    prm = Params(E=E, nu=nu, rho0=rho0, wdeltap=0.0, corr_exp=0.0, mghn_eps=0.0, alpha=alpha, beta=beta,
                 eta=eta, xsph_eps=0.0, dt=dt, time=0.0)
    return PlateScene(ps, prm, left[:nl.value], right[:nr.value], v_bc, dx)


def reference_tensile(nbox):
    """The tension drivers' own scene (tensile_tl, tensile_corot,
    tensile_potential: src/benchmarks/tensile.cpp:137-361): a unit square of
    nbox x nbox particles at (i dx, j dx), dx = 1/(nbox - 1), created with
    idx = i nbox + j, rho0 = 1, h = dx, m = dx^2, E = 1e7, nu = 0.4,
    alpha = beta = 1, eta = 0.1, dt = 1e-6 s, the columns i = 0 and
    i = nbox - 1 pulled apart at 10 m/s."""
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
    left = np.nonzero(ii.ravel() == 0)[0].astype(np.uint32)
    right = np.nonzero(ii.ravel() == nbox - 1)[0].astype(np.uint32)
    prm = Params(E=1e7, nu=0.4, rho0=1.0, wdeltap=0.0, corr_exp=0.0, mghn_eps=0.0, alpha=1.0, beta=1.0, eta=0.1,
                 xsph_eps=0.0, dt=1e-6, time=0.0)
    return PlateScene(ps, prm, left, right, 10.0, dx)


def plate_config3():
    """Tension plate 0.1 x 0.04 m, hole r = 0.01 m, dx = 8e-5 m (~578k particles)."""
    return make_plate()


def splitmix64_stream(seed):
    """SplitMix64 (Steele, Lea, Flood 2014) as a Python generator of uint64."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def batch_specs(nscenes=256, seed=0, nmin=2000, nmax=60000):
    """Per-scene parameters of the batched workload: outer radius U(0.02, 0.05),
    ri/ro U(0.5, 0.85), speed U(20, 100), E log-uniform in [1e6, 1e8]; target
    particle count nmin + (nmax - nmin) * u**1.7 (mean ~23.5k, ~6M in total),
    dx chosen to match it."""
    g = splitmix64_stream(seed)

    def u():
        return ((next(g) >> 11) + 0.5) / 9007199254740992.0

    out = []
    for _ in range(nscenes):
        ro = 0.02 + 0.03 * u()
        ratio = 0.5 + 0.35 * u()
        speed = 20.0 + 80.0 * u()
        E = 10.0 ** (6.0 + 2.0 * u())
        target = nmin + (nmax - nmin) * u() ** 1.7
        ri = ratio * ro
        dx = math.sqrt(2 * math.pi * (ro * ro - ri * ri) / target)
        nbox = max(int(round(2 * ro / dx)) + 1, 8)
        out.append(dict(ri=ri, ro=ro, centre=1.125 * ro, v0=speed, E=E, nbox=nbox))
    return out


def subsonic_batch_specs(nscenes=256, seed=0, nmin=2000, nmax=60000, max_mach=0.5, nu=0.4, rho0=1010.0):
    """The batched workload's generator restricted to impacts below max_mach
    (v0 / c0, c0 = sqrt(K / rho0)): the first nscenes such scenes of the
    seeded stream.  Faster impacts compress the rings far beyond an elastic
    regime (rho/rho0 of order 100 within a few hundred steps) and end in
    non-finite states; the learned stepper is trained on the rest."""
    out, drawn = [], 0
    while len(out) < nscenes:
        batch = batch_specs(nscenes=drawn + 4 * nscenes, seed=seed, nmin=nmin, nmax=nmax)
        out = [b for b in batch if b["v0"] < max_mach * sound_speed(b["E"], nu, rho0)][:nscenes]
        drawn += 4 * nscenes
    return out


def scenes_from_specs(specs):
    return [make_rings(s["nbox"], ri=s["ri"], ro=s["ro"], centre=s["centre"], v0=s["v0"], E=s["E"]) for s in specs]


def batch_config4(nscenes=256, seed=0):
    return [make_rings(s["nbox"], ri=s["ri"], ro=s["ro"], centre=s["centre"], v0=s["v0"], E=s["E"])
            for s in batch_specs(nscenes, seed)]
