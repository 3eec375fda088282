"""Few-particle configurations whose derivatives are known in closed form,
with the expected values written as literals derived by hand from the
published formulas (Monaghan's artificial viscosity and artificial stress,
XSPH), not from the oracle's code:

* artificial viscosity, src/correctors.h:83-137 (UL) and the same corrector
  in the TL action list (src/benchmarks/tensile.cpp:276-286):
      mu_ij = h_ij (v_ij . x_ij) / (r_ij^2 + eta^2 h_ij^2),
      Pi_ij = (-alpha c_ij mu_ij + beta mu_ij^2) / rho_ij    (approaching pairs)
      dv_i/dt -= m_j Pi_ij grad_i W_ij
  with h_ij, c_ij, rho_ij the means of the pair and c = sqrt(K/rho);
* the artificial-stress pair term, src/derivatives.h:424-430:
      f = (W_ij / W(dx))^n,  term f (R_i + R_j) next to sigma_i/rho_i^2 + sigma_j/rho_j^2;
* XSPH, src/correctors.h:139-169: dx_i/dt += -eps W_ij m_i / rho_ij (v_i - v_j)
  (the mass of particle i, src/correctors.h:163).

Units are chosen so the numbers stay simple: E = 3, nu = 0.25 give G = 1.2,
K = 2; rho0 = 2 gives c0 = sqrt(K/rho0) = 1; h = 1, so the cubic spline's
factor is fac = 10/(7 pi) and, for a pair at distance r along x with particle
i on the left, grad_i W = (x_i - x_j)/r dW/dr = (-dW/dr, 0):
    r = 1.5 (q = 1.5):  W = fac (2-q)^3/4 = fac/32,  -dW/dr = 3/4 (2-q)^2 fac = 0.1875 fac
    r = 1.0 (q = 1.0):  W = fac/4,                   -dW/dr = 0.75 fac
Every case below was shown to fail when the restatement is altered in the
way its comment names (the alteration made once in oracle/oracle.cpp, the
test run, the change reverted).
"""
import math

import numpy as np
import pytest

import _oracle as O
from mfree_iwf_amd.scenes import ParticleSet
from mfree_iwf_amd._capi import Params

FAC = 10.0 / (7.0 * math.pi)
GRAD_15 = 0.1875 * FAC  # -dW/dr at q = 1.5, h = 1
GRAD_10 = 0.75 * FAC    # -dW/dr at q = 1.0, h = 1
RTOL = 1e-12


def _params(alpha=1.0, beta=1.0, eta=0.1, art_eps=0.0, wdeltap=0.0, corr_exp=4.0, xsph_eps=0.0):
    return Params(E=3.0, nu=0.25, rho0=2.0, wdeltap=wdeltap, corr_exp=corr_exp, mghn_eps=art_eps, alpha=alpha,
                  beta=beta, eta=eta, xsph_eps=xsph_eps, dt=1e-3, time=0.0)


def _set(pos, m, rho, h, v, S=None):
    n = len(pos)
    ps = ParticleSet.zeros(n)
    ps.x[:] = [p[0] for p in pos]
    ps.y[:] = [p[1] for p in pos]
    ps.X[:] = ps.x
    ps.Y[:] = ps.y
    ps.m[:] = m
    ps.rho[:] = rho
    ps.h[:] = h
    ps.vx[:] = [u[0] for u in v]
    ps.vy[:] = [u[1] for u in v]
    if S is not None:
        ps.Sxx[:] = [s[0] for s in S]
        ps.Sxy[:] = [s[1] for s in S]
        ps.Syy[:] = [s[2] for s in S]
    return ps


def _ul(ps, prm):
    rc, d8, aux = O.ul_derivatives(ps, prm)
    assert rc == O.OK
    return d8


# --- artificial viscosity (UL) ---------------------------------------------
# Particles 0 at (0, 0) and 1 at (1.5, 0), v0 = (2, 0), v1 = (-2, 0), rho = rho0
# (p = 0, S = 0: no stress force, R = 0), m1 = 0.75, m0 = 0.5:
#   v_ij . x_ij = (2 - (-2)) (0 - 1.5) = -6 < 0 (approaching)
#   particle 0: vx_t = -m1 Pi GRAD_15;  particle 1: vx_t = +m0 Pi GRAD_15
# V1  alpha = 1, beta = 0, eta = 0.1: mu = -6/(2.25 + 0.01) = -6/2.26,
#     Pi = c mu... = 1 * (6/2.26) / 2 = 3/2.26            (a wrong alpha scale fails)
# V2  alpha = 0, beta = 1: Pi = (6/2.26)^2 / 2            (a dropped beta mu^2 fails)
# V3  alpha = beta = 1, eta = 0.5: eta^2 h^2 = 0.25, mu = -6/2.5 = -2.4,
#     Pi = (2.4 + 5.76)/2 = 4.08                           (eta in place of eta^2 fails)
VISC_CASES = {
    "V1": (dict(alpha=1.0, beta=0.0, eta=0.1), 3.0 / 2.26),
    "V2": (dict(alpha=0.0, beta=1.0, eta=0.1), (6.0 / 2.26) ** 2 / 2.0),
    "V3": (dict(alpha=1.0, beta=1.0, eta=0.5), 4.08),
}


@pytest.mark.parametrize("case", sorted(VISC_CASES))
def test_ul_viscosity_head_on_pair(case):
    kw, pi_ij = VISC_CASES[case]
    ps = _set([(0.0, 0.0), (1.5, 0.0)], m=[0.5, 0.75], rho=2.0, h=1.0, v=[(2.0, 0.0), (-2.0, 0.0)])
    d8 = _ul(ps, _params(**kw))
    # the literals of the hand derivation (V1: -0.75 * 1.3274336283185841 * 0.0852615766563725)
    lit = {"V1": (-0.08488431304284872918, 0.05658954202856581945),
           "V2": (-0.22535658329959839604, 0.15023772219973226403),
           "V3": (-0.26090042456849985399, 0.17393361637899990266)}[case]
    assert d8[0, 2] == pytest.approx(-0.75 * pi_ij * GRAD_15, rel=RTOL)
    assert d8[0, 2] == pytest.approx(lit[0], rel=RTOL)
    assert d8[1, 2] == pytest.approx(lit[1], rel=RTOL)
    assert d8[0, 3] == 0.0 and d8[1, 3] == 0.0


def test_ul_viscosity_means_of_unequal_pair():
    # V4: rho1 = 8 (c1 = sqrt(2/8) = 0.5, S1 = p1 I = 6 I so sigma1 = 0), h1 = 1.2;
    # alpha = beta = 1, eta = 0.1.  c_ij = 0.75, rho_ij = 5, h_ij = 1.1:
    #   mu = 1.1 * (-6) / (2.25 + 0.01 * 1.21) = -6.6 / 2.2621
    #   Pi = (0.75 * 6.6/2.2621 + (6.6/2.2621)^2) / 5 = 2.14017445100277092
    # particle 0 (kernel with h0 = 1): vx_t = -0.75 Pi GRAD_15;
    # particle 1 (kernel with h1 = 1.2: q = 1.25, fac1 = fac/1.44,
    #   -dW/dr = 0.75 (0.75)^2 / 1.2 * fac1 = 0.111017677937985): vx_t = +0.5 Pi 0.111...
    # c_ij from sqrt(K/rho_ij) (0.632) or rho_i in place of rho_ij fails.
    ps = _set([(0.0, 0.0), (1.5, 0.0)], m=[0.5, 0.75], rho=[2.0, 8.0], h=[1.0, 1.2],
              v=[(2.0, 0.0), (-2.0, 0.0)], S=[(0.0, 0.0, 0.0), (6.0, 0.0, 6.0)])
    d8 = _ul(ps, _params(alpha=1.0, beta=1.0, eta=0.1))
    assert d8[0, 2] == pytest.approx(-0.13685598600913701529, rel=RTOL)
    assert d8[1, 2] == pytest.approx(0.11879859896626477022, rel=RTOL)


def test_ul_viscosity_receding_and_transverse_pairs_are_zero():
    # receding (v reversed) and transverse (v perpendicular to x_ij: v.x = 0)
    for v in ([(-2.0, 0.0), (2.0, 0.0)], [(0.0, 2.0), (0.0, -2.0)]):
        ps = _set([(0.0, 0.0), (1.5, 0.0)], m=[0.5, 0.75], rho=2.0, h=1.0, v=v)
        d8 = _ul(ps, _params(alpha=1.0, beta=1.0, eta=0.1))
        assert np.all(d8[:, 2:4] == 0.0)


# --- artificial-stress pair term ------------------------------------------
# rho = rho0 = 2 (p = 0, sigma = S), v = 0 (no viscosity, no XSPH, no grad v),
# eps = 0.3, n = 4, wdeltap = W(1.0) = fac/4, m1 = 0.75.
# A1  pair along x at r = 1.5, S0 = diag(4, -1), S1 = diag(2, -3):
#     theta = 0, R0 = (-0.3*4/4, 0, 0) = (-0.3, 0, 0), R1 = (-0.15, 0, 0),
#     f = ((fac/32)/(fac/4))^4 = (1/8)^4 = 1/4096,
#     vx_t0 = (1/rho0) rho0 m1 (4/4 + 2/4 + f (R0 + R1)) GRAD_15
#           = 0.75 (1.5 - 0.45/4096) GRAD_15              (a wrong exponent or a dropped R_j fails)
# A2  the same at r = 1.0: f = 1, vx_t0 = 0.75 (1.5 - 0.45) GRAD_10
# A3  the pair along y, S0 = diag(-1, 4), S1 = diag(-3, 2): the tensile axis is
#     y (theta = pi/2), R0 = (0, 0, -0.3), R1 = (0, 0, -0.15):
#     vy_t0 = 0.75 (1.5 - 0.45/4096) GRAD_15              (R rotated back wrongly fails)
# A4  pure shear S0 = S1 = [[0, 2], [2, 0]] at r = 1.0 along x: theta = pi/4,
#     principal (2, -2) -> R' = (-0.15, 0), R = (-0.075, -0.075, -0.075):
#     vx_t0 = 0.75 (0 + 0 - 0.15) GRAD_10, vy_t0 = 0.75 (0.5 + 0.5 - 0.15) GRAD_10
#                                                          (a wrong sign of R_xy fails)
AS_CASES = {
    "A1": ([(0.0, 0.0), (1.5, 0.0)], [(4.0, 0.0, -1.0), (2.0, 0.0, -3.0)], (0.09591224840098705085, 0.0)),
    "A2": ([(0.0, 0.0), (1.0, 0.0)], [(4.0, 0.0, -1.0), (2.0, 0.0, -3.0)], (0.26857396646757337911, 0.0)),
    "A3": ([(0.0, 0.0), (0.0, 1.5)], [(-1.0, 0.0, 4.0), (-3.0, 0.0, 2.0)], (0.0, 0.09591224840098705085)),
    "A4": ([(0.0, 0.0), (1.0, 0.0)], [(0.0, 2.0, 0.0), (0.0, 2.0, 0.0)],
           (-0.03836770949536762559, 0.21741702047374987833)),
}


@pytest.mark.parametrize("case", sorted(AS_CASES))
def test_artificial_stress_pair_term(case):
    pos, S, (ax, ay) = AS_CASES[case]
    ps = _set(pos, m=[0.5, 0.75], rho=2.0, h=1.0, v=[(0.0, 0.0), (0.0, 0.0)], S=S)
    prm = _params(alpha=1.0, beta=1.0, eta=0.1, art_eps=0.3, wdeltap=FAC / 4.0, corr_exp=4.0)
    d8 = _ul(ps, prm)
    scale = abs(ax) + abs(ay)
    assert d8[0, 2] == pytest.approx(ax, rel=RTOL, abs=1e-14 * scale)
    assert d8[0, 3] == pytest.approx(ay, rel=RTOL, abs=1e-14 * scale)


def test_artificial_stress_exponent_other_than_four():
    # n = 3 at r = 1.5: f = (1/8)^3 = 1/512; vx_t0 = 0.75 (1.5 - 0.45/512) GRAD_15
    # = 0.09586307103896294 (a hard-wired fourth power fails)
    pos, S, _ = AS_CASES["A1"]
    ps = _set(pos, m=[0.5, 0.75], rho=2.0, h=1.0, v=[(0.0, 0.0), (0.0, 0.0)], S=S)
    prm = _params(art_eps=0.3, wdeltap=FAC / 4.0, corr_exp=3.0)
    d8 = _ul(ps, prm)
    assert d8[0, 2] == pytest.approx(0.75 * (1.5 - 0.45 / 512.0) * GRAD_15, rel=RTOL)
    assert d8[0, 2] == pytest.approx(0.09586307103896294, rel=RTOL)


# --- XSPH ----------------------------------------------------------------
def test_xsph_unequal_masses():
    # X1: m0 = 0.5, m1 = 1.5, rho = (2, 8) (rho_ij = 5), h = 1, r = 1.5 along x,
    # v0 = (1, 3), v1 = (-2, 0.5), xsph eps = 0.5, W = fac/32:
    #   x_t0 = 1 - 0.5 (fac/32) 0.5/5 (1 - (-2)) = 0.99786846058359068747
    #   y_t0 = 3 - 0.5 (fac/32) 0.5/5 (3 - 0.5)  = 2.99822371715299223956
    #   x_t1 = -2 - 0.5 (fac/32) 1.5/5 (-2 - 1)   = -1.99360538175077206240
    #   y_t1 = 0.5 - 0.5 (fac/32) 1.5/5 (0.5 - 3) = 0.50532884854102328133
    # m_j in place of m_i, or eps W m / rho_i, fails.
    ps = _set([(0.0, 0.0), (1.5, 0.0)], m=[0.5, 1.5], rho=[2.0, 8.0], h=1.0, v=[(1.0, 3.0), (-2.0, 0.5)],
              S=[(0.0, 0.0, 0.0), (6.0, 0.0, 6.0)])
    d8 = _ul(ps, _params(xsph_eps=0.5))
    assert d8[0, 0] == pytest.approx(0.99786846058359068747, rel=RTOL)
    assert d8[0, 1] == pytest.approx(2.99822371715299223956, rel=RTOL)
    assert d8[1, 0] == pytest.approx(-1.99360538175077206240, rel=RTOL)
    assert d8[1, 1] == pytest.approx(0.50532884854102328133, rel=RTOL)


# --- artificial viscosity (TL) --------------------------------------------
def test_tl_viscosity_plus_stencil():
    # Five particles in a plus, spacing 1, h = 1, m = 0.5, rho = rho0 = 2
    # (V = 0.25, c = 1), at their reference positions (F = I, sigma = 0, P = 0).
    # Centre particle: C = V (W(0) + 4 W(1)); the moment matrix
    # B = sum_j (x_j - x_i) (x) grad W_ij V_j / C = 1.5 fac V / C * I, so the
    # corrected gradient toward (1, 0) is grad W / (1.5 fac V) = (0.75 fac)/(0.375 fac)
    # = (2, 0) = (0.5/V, 0).
    # The centre moves with v = (3, 0), the others rest: only the pair with
    # (1, 0) approaches (v.x = -3); (0, +-1) have v.x = 0 and (-1, 0) recedes.
    #   mu = -3/(1 + 0.01) = -2.9702970297029703,
    #   Pi = (2.9702970297029703 + 8.822664444662287)/2 = 5.896480737182628,
    #   vx_t(centre) = -m Pi 2 = -Pi.
    ps = _set([(0.0, 0.0), (1.0, 0.0), (-1.0, 0.0), (0.0, 1.0), (0.0, -1.0)], m=0.5, rho=2.0, h=1.0,
              v=[(3.0, 0.0)] + [(0.0, 0.0)] * 4)
    t = O.TL(ps, 1.0, search=1, contrib_only=1)
    assert t.rc == O.OK
    try:
        row, col, w = t.lists()
        js = list(col[row[0]:row[1]])
        k = js.index(1)
        assert w[row[0] + k, 1] == pytest.approx(2.0, rel=1e-13)  # the corrected gradient derived above
        rc, d8, aux, _ = t.derivatives(ps, _params(alpha=1.0, beta=1.0, eta=0.1))
        assert rc == O.OK
        assert d8[0, 2] == pytest.approx(-5.896480737182628, rel=RTOL)
        assert d8[0, 3] == pytest.approx(0.0, abs=1e-15)
        for i in (2, 3, 4):
            assert d8[i, 2] == 0.0 and d8[i, 3] == 0.0
        # beta = 0: Pi = 2.9702970297029703 / 2 -> vx_t = -1.4851485148514851
        rc, d8, aux, _ = t.derivatives(ps, _params(alpha=1.0, beta=0.0, eta=0.1))
        assert d8[0, 2] == pytest.approx(-1.4851485148514851, rel=RTOL)
    finally:
        t.close()
