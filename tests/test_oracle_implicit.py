"""CPU pins of the oracle's implicit-Euler dual-time step (ImplicitEulerDualTimeStep::Step,
src/Solver/ODE.hpp:80-98, with EulerSolver's fdtau / frhs / fsolve / fincrement / fstop).

Pinned by what the mathematics fixes rather than by a reference number (none exists for
this loop): the converged state solves the discrete implicit-Euler equation
(u - xLast) / dt = R(u), whatever the preconditioner, so Jacobi and Gauss-Seidel sweeps,
and any number of extra sweeps, must reach the same state; and the first pseudo-time
iteration from a zero increment equals one LUSGSMatrixInit + symmetric sweep computed
from the oracle's own operators."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import fields
from dndsr_amd import mesh as mm
from helpers import RS, rel_l2


def _iv_case(n=10, order=1, rs="HLLEP"):
    m = mm.periodic_quad_grid(n, n, 10.0, 10.0, jitter=0.1, seed=3)
    o = po.OracleVR(m, maxOrder=order, intOrder=3 if order == 1 else 5)
    bg = fields.iv_background_for_mach(0.5, 1.4)
    u = o.cell_average(10, [5.0, 1.4, bg[0], bg[1]], 5)
    es = o.euler_settings(1.4, RS[rs])
    return o, es, u


def _explicit_dt(o, es, u, CFL=0.5):
    _, _, _, dtMin = o.eval_dt(es, u, CFL, 1e100, False)
    return dtMin


def _implicit_residual(o, es, u, uRec, xLast, dt):
    _, _, dlc, _ = o.eval_dt(es, u, 1.0, 1e100, True)
    r = o.eval_rhs(es, u, uRec, dlc)
    return (xLast - u) / dt + r


@pytest.mark.parametrize("rs", ["HLLEP", "Roe"])
def test_converged_state_solves_implicit_euler(rs):
    o, es, u0 = _iv_case(rs=rs)
    dt = 8 * _explicit_dt(o, es, u0)
    rec0 = np.zeros((o.nCell, 5, o.M))
    u, rec, n, res, rel, hist = o.implicit_euler_step(es, u0, rec0, dt, CFL=10.0, maxIter=400, minIter=5,
                                                      tol=1e-12)
    assert n < 400 and rel[0] < 1e-12
    # the density residual relative to its running maximum is what the stop test sees
    assert hist[-1] == rel[0] and np.all(np.diff(np.maximum.accumulate(hist)) <= 0)
    # the reconstruction has converged with the state
    assert rel_l2(o.vr_iter(rec, u), rec) < 1e-9
    r = _implicit_residual(o, es, u, rec, u0, dt)
    r0 = o.eval_rhs(es, u0, o.vr_iter(rec0, u0), o.eval_dt(es, u0, 1.0, 1e100, True)[2])
    assert np.abs(r).sum() < 1e-9 * np.abs(r0).sum()
    assert rel_l2(u, u0) > 1e-4  # the step moved the state


def test_preconditioner_does_not_change_the_converged_state():
    o, es, u0 = _iv_case()
    dt = 8 * _explicit_dt(o, es, u0)
    rec0 = np.zeros((o.nCell, 5, o.M))
    ug, *_ = o.implicit_euler_step(es, u0, rec0, dt, gsUpdate=1, maxIter=400, tol=1e-13)
    uj, *_ = o.implicit_euler_step(es, u0, rec0, dt, gsUpdate=0, CFL=5.0, maxIter=800, tol=1e-13)
    us, *_ = o.implicit_euler_step(es, u0, rec0, dt, gsUpdate=1, sgsIter=2, maxIter=400, tol=1e-13)
    assert rel_l2(uj, ug) < 1e-10
    assert rel_l2(us, ug) < 1e-10


def test_first_iteration_matches_operator_sequence():
    """maxIter = 1: dTau from EvaluateDt, one VR sweep, RHS, LUSGSMatrixInit, forward then
    backward sweep from a zero increment, u += inc."""
    o, es, u0 = _iv_case(n=6)
    dt = 4 * _explicit_dt(o, es, u0)
    rec0 = np.random.default_rng(0).uniform(-1e-3, 1e-3, (o.nCell, 5, o.M))
    order = np.random.default_rng(1).permutation(o.nCell).astype(np.int64)
    u1, rec1, n, *_ = o.implicit_euler_step(es, u0, rec0, dt, CFL=10.0, maxIter=1, minIter=1, scanOrder=order)
    assert n == 1
    dTau, lam, dlc, _ = o.eval_dt(es, u0, 10.0, 1e100, True)
    rec = o.vr_iter(rec0, u0)
    rhs = o.eval_rhs(es, u0, rec, dlc)  # xLast == u on the first iteration
    jd = o.lusgs_init(dTau, dt, 1.0, lam)
    z = np.zeros_like(u0)
    x1, _ = o.update_sgs(es, 1.0, rhs, u0, z, z, jd, lam, True, True, False, scanOrder=order)
    x2, _ = o.update_sgs(es, 1.0, rhs, u0, x1, z, jd, lam, False, True, False, scanOrder=order)
    np.testing.assert_array_equal(rec1, rec)
    np.testing.assert_allclose(u1, u0 + x2, rtol=0, atol=1e-14 * np.abs(u0).max())


def test_sgs_converges_to_the_lusgs_matrix_solution():
    """LUSGSMatrixVec is the operator UpdateSGS inverts: repeated symmetric Gauss-Seidel
    sweeps converge to x with lusgs_matvec(x) = b (pins the oracle's matvec)."""
    from helpers import random_state
    m = mm.periodic_quad_grid(6, 5, 6.0, 5.0, jitter=0.1, seed=5)
    o = po.OracleVR(m, maxOrder=1, intOrder=1)
    es = o.euler_settings(1.4, 1)
    N = o.nCell
    u = random_state(N, 6)
    dt, lam, _, _ = o.eval_dt(es, u, 1.0)
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    b = np.random.default_rng(3).uniform(-1, 1, (N, 5))
    x = np.zeros((N, 5))
    z = np.zeros_like(x)
    for _ in range(200):
        x1, _ = o.update_sgs(es, 1.0, b, u, x, z, jd, lam, True, True, False)
        x, _ = o.update_sgs(es, 1.0, b, u, x1, z, jd, lam, False, True, False)
    assert np.abs(o.lusgs_matvec(es, 1.0, u, x, jd, lam) - b).max() < 1e-13
