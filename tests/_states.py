"""Seeded non-trivial particle states shared by the oracle and GPU tests."""
import numpy as np

from mfree_iwf_amd import scenes


def perturbed_rings(nbox=41, seed=0, dv=2.0, drho=0.01, dS=2e5, dpos=0.1, random_mass=False, shuffle=False,
                    **kw):
    """Ring scene with every field perturbed (positions by dpos*dx)."""
    s = scenes.make_rings(nbox, shuffle=shuffle, seed=seed, **kw)
    ps = s.particles
    rng = np.random.default_rng(seed + 1000)
    n = ps.n
    ps.x += dpos * s.dx * rng.uniform(-1, 1, n)
    ps.y += dpos * s.dx * rng.uniform(-1, 1, n)
    ps.vx += dv * rng.uniform(-1, 1, n)
    ps.vy += dv * rng.uniform(-1, 1, n)
    ps.rho *= 1 + drho * rng.uniform(-1, 1, n)
    ps.Sxx += dS * rng.uniform(-1, 1, n)
    ps.Sxy += dS * rng.uniform(-1, 1, n)
    ps.Syy += dS * rng.uniform(-1, 1, n)
    if random_mass:
        ps.m *= rng.uniform(0.5, 1.5, n)
    return s


def rotate90(ps):
    """(x, y) -> (-y, x) with vectors and tensors rotated alike."""
    q = ps.copy()
    q.x, q.y = -ps.y.copy(), ps.x.copy()
    q.X, q.Y = -ps.Y.copy(), ps.X.copy()
    q.vx, q.vy = -ps.vy.copy(), ps.vx.copy()
    q.Sxx, q.Syy, q.Sxy = ps.Syy.copy(), ps.Sxx.copy(), -ps.Sxy.copy()
    return q


def rotate90_d8(d):
    """Derivatives of the rotated state expressed from the original ones."""
    r = np.empty_like(d)
    r[:, 0], r[:, 1] = -d[:, 1], d[:, 0]
    r[:, 2], r[:, 3] = -d[:, 3], d[:, 2]
    r[:, 4] = d[:, 4]
    r[:, 5], r[:, 7], r[:, 6] = d[:, 7], d[:, 5], -d[:, 6]
    return r
