"""Pins the oracle's positivity-preserving increment (CPU only):
EvaluateCellRHSAlpha (src/Euler/EulerEvaluator.hxx:1938-2030) with
IdealGasGetCompressionRatioPressure (src/Euler/Gas.hpp:1750-1860), CompressInc
(src/Euler/EulerEvaluator.hpp:1645-1760) and their use in the implicit step's fincrement
(src/Euler/EulerSolver.hxx:696-738).

The limiter's targets are fixed by the mathematics: the density limit scales the increment so
that rho lands on the relaxed bound; the pressure limit is the root of the internal energy
rho e(u + a d) = target, which is a quadratic in a and is found here by bisection on the
textbook internal energy, independent of the quadratic's coefficients."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import mesh as mm
from helpers import RS

G = 1.4
SAFE = 1 - 1e-5
RHO_EPS = P_EPS = 1e-11   # smallReal * refRho * 0.1 with refRho = refP = 1


def rho_e(U):
    return U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]


def pressure(U):
    return (G - 1) * rho_e(U)


def bisect(g, lo=0.0, hi=1.0, n=200):
    """Root of g on [lo, hi] with g(lo) > 0 > g(hi)."""
    for _ in range(n):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


U0 = np.array([1.0, 0.3, -0.2, 0.1, 2.6])


def test_small_increment_is_untouched():
    d = np.array([0.01, -0.02, 0.01, 0.0, -0.03])
    a, lim = po.cell_rhs_alpha(U0, d, G, 0.9, RHO_EPS, P_EPS)
    assert a == 1.0 and not lim
    np.testing.assert_array_equal(po.compress_inc(U0, d, G, RHO_EPS, P_EPS), d)


@pytest.mark.parametrize("relax", [0.9, 0.5, 1.0])
def test_density_limit_lands_on_relaxed_bound(relax):
    """alpha_rho puts rho on rhoEps relax + rho (1 - relax) (up to the smallReal guard in the
    denominator); the returned alpha is alpha_rho^2 (compress 2) times the safety ratio."""
    U = np.array([1.0, 0.0, 0.0, 0.0, 2.6])   # at rest: removing mass keeps the pressure
    d = np.array([-3.0, 0.0, 0.0, 0.0, 0.0])
    a, lim = po.cell_rhs_alpha(U, d, G, relax, RHO_EPS, P_EPS)
    assert lim
    target = RHO_EPS * relax + U[0] * (1 - relax)
    ar = np.sqrt(a / SAFE)
    assert abs((U[0] + ar * d[0]) - target) <= 1e-9 * U[0]
    assert pressure(U + ar * d) == pytest.approx(pressure(U), rel=1e-14)


def test_pressure_limit_is_the_internal_energy_root():
    """A pure energy decrement: rho e is linear in alpha, the root is (rho e - target) / x."""
    x = 3.0
    d = np.array([0.0, 0.0, 0.0, 0.0, -x])
    relax = 0.9
    a, lim = po.cell_rhs_alpha(U0, d, G, relax, RHO_EPS, P_EPS)
    assert lim
    pOld = pressure(U0)
    target = (P_EPS + (pOld - P_EPS) * (1 - relax)) / (G - 1)
    root = (rho_e(U0) - target) / x
    np.testing.assert_allclose(a, (root * SAFE) ** 2 * SAFE, rtol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_pressure_limit_quadratic_case(seed):
    """General increments (momentum and density change too): rho e(u + a d) is a ratio of
    quadratics in a; the limiter's alpha is the bisection root of rho e(u + a d) = target
    (times the safety ratio, then squared)."""
    rng = np.random.default_rng(seed)
    d = np.concatenate([[rng.uniform(-0.3, 0.3)], rng.uniform(-1.0, 1.0, 3), [-rng.uniform(2.5, 4.0)]])
    relax = 0.9
    a, lim = po.cell_rhs_alpha(U0, d, G, relax, RHO_EPS, P_EPS)
    # the density part first (alpha_rho), then the pressure root on the scaled increment
    ar = 1.0
    if d[0] < 0:
        ar = min(1.0, (U0[0] - (RHO_EPS * relax + U0[0] * (1 - relax))) / (-d[0] - 1e-10 * d[0]))
    dd = ar * d
    pOld = pressure(U0)
    if pressure(U0 + dd) >= P_EPS + (pOld - P_EPS) * (1 - relax):
        pytest.skip("this draw needs no pressure limit")
    target = (P_EPS + (pOld - P_EPS) * (1 - relax)) / (G - 1)
    root = bisect(lambda s: rho_e(U0 + s * dd) - target)
    assert lim
    np.testing.assert_allclose(a, (ar * root * SAFE) ** 2 * SAFE, rtol=1e-9)


def test_compress_inc_density_branch():
    """rho + d_rho <= rhoEps: the whole increment is scaled so that rho lands on rhoEps (a state
    and increment without momentum, so that the internal energy stays positive)."""
    U = np.array([1.0, 0.0, 0.0, 0.0, 2.6])
    d = np.array([-2.0, 0.0, 0.0, 0.0, -0.5])
    r = po.compress_inc(U, d, G, RHO_EPS, P_EPS)
    s = r[0] / d[0]
    np.testing.assert_allclose(r, s * d, rtol=1e-14)
    assert abs(U[0] + r[0] - RHO_EPS) <= 1e-15


def test_compress_inc_energy_branch():
    """rho e(u + d) <= pEps: the increment is scaled by at most 0.99 of the internal-energy
    root and leaves rho e above pEps / (gamma - 1)."""
    d = np.array([0.2, -0.4, 0.3, 0.1, -4.0])
    assert rho_e(U0 + d) <= P_EPS
    r = po.compress_inc(U0, d, G, RHO_EPS, P_EPS)
    s = r[4] / d[4]
    np.testing.assert_allclose(r, s * d, rtol=1e-13)
    root = bisect(lambda t: rho_e(U0 + t * d) - P_EPS / (G - 1))
    assert s <= 0.99 * root * (1 + 1e-12)
    assert rho_e(U0 + r) >= P_EPS / (G - 1)
    assert s >= 0.99 * root * 0.9 ** 2   # at most a couple of decay steps


def test_fincrement_composes_alpha_and_compression():
    """u_i + CompressInc(u_i, alpha_i d_i) per cell, with alpha_i from EvaluateCellRHSAlpha."""
    rng = np.random.default_rng(3)
    N = 40
    U = np.tile(U0, (N, 1)) * rng.uniform(0.8, 1.2, (N, 1))
    D = rng.normal(size=(N, 5)) * np.array([0.8, 1, 1, 1, 3])
    out, nLim, aMin = po.add_increment_pp(U, D, G, 0.9, RHO_EPS, P_EPS)
    ref = U.copy()
    cnt, amin = 0, 1.0
    for i in range(N):
        a, lim = po.cell_rhs_alpha(U[i], D[i], G, 0.9, RHO_EPS, P_EPS)
        cnt += lim
        if lim:
            amin = min(amin, a)
        ref[i] += po.compress_inc(U[i], a * D[i], G, RHO_EPS, P_EPS)
    np.testing.assert_array_equal(out, ref)
    assert nLim == cnt > 0 and aMin == amin
    assert (out[:, 0] > 0).all() and all(rho_e(x) > 0 for x in out)


def blast_case(n=12):
    """A first-order blast (density 20x, pressure 1000x inside a disc) on a jittered periodic
    quad mesh: at pseudo-CFL 5 the implicit increments need the positivity limiter."""
    m = mm.periodic_quad_grid(n, 10, 1.2, 1.0, jitter=0.1, seed=4)
    o = po.OracleVR(m, maxOrder=1, intOrder=3)
    bary = o.cell_metrics()["bary"]
    inside = np.linalg.norm(bary[:, :2] - [0.6, 0.5], axis=1) < 0.2
    r = np.where(inside, 1.0, 0.05)
    p = np.where(inside, 10.0, 0.01)
    u = np.stack([r, 0 * r, 0 * r, 0 * r, p / (G - 1)], 1)
    return m, o, u


BLAST = dict(dt=1e100, CFL=5.0, nRec=0, sgsIter=1, maxIter=5, minIter=5, tol=1e-30, incrementRelax=0.9,
             refRho=1.0, refP=1.0)


def test_implicit_step_applies_the_limiter():
    m, o, u = blast_case()
    es = o.euler_settings(G, RS["Roe"])
    out = o.implicit_euler_step(es, u, np.zeros((o.nCell, 5, o.M)), stats=True, **BLAST)
    un, nLim, aMin = out[0], out[-1][0], out[-1][1]
    assert nLim > 0 and 0 < aMin < 1
    assert (un[:, 0] > 0).all() and all(rho_e(x) > 0 for x in un)
