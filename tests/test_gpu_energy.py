"""GPU checks of the FP64 energy diagnostic (mfree_energy) against closed
forms and an independent quadrature of the EOS work, and the energy
trajectory of the mixed mode against the FP64 mode."""
import numpy as np
import pytest
from scipy.integrate import quad

from mfree_iwf_amd import PRECISION_FP64, PRECISION_MIXED, Solver, scenes

pytestmark = pytest.mark.gpu


def _solver(scene, precision=PRECISION_FP64):
    sv = Solver(precision=precision)
    sv.add_body(scene.particles, scene.params)
    return sv


def _uniform(scene, vx=0.0, vy=0.0, rho_scale=1.0, sxy=0.0):
    p = scene.particles.copy()
    n = p.n
    p.vx = np.full(n, vx)
    p.vy = np.full(n, vy)
    p.rho = np.full(n, scene.params.rho0 * rho_scale)
    p.Sxx = np.zeros(n)
    p.Syy = np.zeros(n)
    p.Sxy = np.full(n, sxy)
    return p


def _energy_of(scene, p):
    with Solver(precision=PRECISION_FP64) as sv:
        sv.add_body(p, scene.params)
        return np.array(sv.energy(0))


def test_energy_closed_forms():
    sc = scenes.rings_config1()
    prm = sc.params
    m = sc.particles.m
    print(f"[energy closed forms] n={sc.particles.n}", flush=True)
    # rest state: every term is exactly zero
    assert np.array_equal(_energy_of(sc, _uniform(sc)), np.zeros(3))
    # rigid translation: kinetic = M |u|^2 / 2, nothing else
    e = _energy_of(sc, _uniform(sc, vx=3.0, vy=-4.0))
    assert e[0] == pytest.approx(0.5 * m.sum() * 25.0, rel=1e-13)
    assert e[1] == 0.0 and e[2] == 0.0
    # pure shear stress: (m/rho) * 2 tau^2 / (4 G), G = E / (2 (1 + nu))
    tau = 2.5e5
    G = prm.E / (2.0 * (1.0 + prm.nu))
    e = _energy_of(sc, _uniform(sc, sxy=tau))
    assert e[2] == pytest.approx((m / prm.rho0).sum() * 2.0 * tau * tau / (4.0 * G), rel=1e-13)
    assert e[0] == 0.0 and e[1] == 0.0


@pytest.mark.parametrize("d", [1e-3, 2e-2, -5e-2, -3e-4])
def test_energy_volumetric_is_eos_work(d):
    """Uniform compression: m * integral_{rho0}^{rho} p(r) / r^2 dr with the
    EOS p = c0^2 (r - rho0), by adaptive quadrature."""
    sc = scenes.rings_config1()
    prm = sc.params
    c0sq = scenes.elastic_K(prm.E, prm.nu) / prm.rho0
    rho = prm.rho0 * (1.0 + d)
    work, _ = quad(lambda r: c0sq * (r - prm.rho0) / (r * r), prm.rho0, rho, epsabs=0.0, epsrel=1e-13)
    e = _energy_of(sc, _uniform(sc, rho_scale=1.0 + d))
    print(f"[energy volumetric] d={d} e={e[1]:.6e} quad={work * sc.particles.m.sum():.6e}", flush=True)
    assert e[1] == pytest.approx(work * sc.particles.m.sum(), rel=1e-9)
    assert e[1] > 0.0


def test_energy_trajectory_mixed_vs_fp64():
    """The ring collision dissipates through artificial viscosity; the mixed
    mode follows the FP64 energy trajectory."""
    sc = scenes.rings_config1()
    out = {}
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        with _solver(sc, prec) as sv:
            e0 = sum(sv.energy(0))
            traj = [e0]
            for _ in range(4):
                sv.step(50)
                traj.append(sum(sv.energy(0)))
            out[name] = np.array(traj)
    print(f"[energy trajectory] fp64 {out['fp64'] / out['fp64'][0]}", flush=True)
    assert np.all(out["fp64"][1:] <= out["fp64"][0] * (1.0 + 1e-6))
    assert np.max(np.abs(out["mixed"] - out["fp64"]) / out["fp64"][0]) < 1e-6
