"""CPU pins of the oracle's BDF dual-time step (ImplicitBDFDualTimeStep::Step,
src/Solver/ODE.hpp:427-477; coefficients :590-595).

Pinned by what the mathematics fixes: BDF1 is the implicit-Euler step (to the last bit but the contraction of a multiply-add); a
converged BDF2 step solves its discrete equation (3 x - 4 u^n + u^{n-1}) / (2 dt) = R(x); and
on a smooth unsteady flow the BDF2 march converges at second order in dt, which a wrong
coefficient would break."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import fields
from dndsr_amd import mesh as mm
from helpers import RS, rel_l2


def _iv(n=8):
    m = mm.periodic_quad_grid(n, n, 10.0, 10.0, jitter=0.1, seed=3)
    o = po.OracleVR(m, maxOrder=1, intOrder=3)
    bg = fields.iv_background_for_mach(0.5, 1.4)
    u = o.cell_average(10, [5.0, 1.4, bg[0], bg[1]], 5)
    es = o.euler_settings(1.4, RS["Roe"])
    return o, es, u


def _march(o, es, u0, dt, nStep, k, tol=1e-12):
    u, rec = u0.copy(), np.zeros((o.nCell, 5, o.M))
    prevs = []
    for _ in range(nStep):
        un, rec, n, _, rel, _ = o.implicit_bdf_step(es, u, rec, dt, prevs[: k - 1], CFL=20.0, sgsIter=1,
                                                    maxIter=500, minIter=3, tol=tol)
        assert rel[0] < tol, (n, rel)
        prevs = [u] + prevs
        u = un
    return u


def test_bdf1_is_implicit_euler():
    o, es, u0 = _iv()
    dt = 0.3
    rec0 = np.zeros((o.nCell, 5, o.M))
    a = o.implicit_euler_step(es, u0, rec0, dt, CFL=20.0, maxIter=6, minIter=6, tol=0.0)
    b = o.implicit_bdf_step(es, u0, rec0, dt, [], CFL=20.0, maxIter=6, minIter=6, tol=0.0)
    assert rel_l2(b[0], a[0]) < 1e-15 and rel_l2(b[1], a[1]) < 1e-14


def test_converged_bdf2_step_solves_its_equation():
    o, es, u0 = _iv()
    dt = 0.3
    u1 = _march(o, es, u0, dt, 1, 1)
    u2, rec, n, _, rel, _ = o.implicit_bdf_step(es, u1, np.zeros((o.nCell, 5, o.M)), dt, [u0], CFL=20.0,
                                                sgsIter=1, maxIter=500, minIter=3, tol=1e-13)
    _, _, dlc, _ = o.eval_dt(es, u2, 1.0, 1e100, True)
    R = o.eval_rhs(es, u2, rec, dlc)
    r = (3 * u2 - 4 * u1 + u0) / (2 * dt) - R
    assert np.abs(r).sum() < 1e-9 * np.abs(R).sum()
    assert rel_l2(u2, u1) > 1e-4


def test_bdf_temporal_order():
    """Errors at t = 1.2 against a fine-step march of the same scheme.  BDF2 converges at second
    order.  BDF3 starts, as the reference's stepper does, with a BDF1 and a BDF2 step, whose
    start-up error keeps its global order at two; its coefficients still make it consistent and
    more accurate than BDF2 at every step size."""
    o, es, u0 = _iv()
    T = 1.2
    errs = {}
    for k in (2, 3):
        ref = _march(o, es, u0, T / 64, 64, k)
        errs[k] = [rel_l2(_march(o, es, u0, T / nS, nS, k), ref) for nS in (4, 8, 16)]
    rates = [np.log2(errs[2][i] / errs[2][i + 1]) for i in range(2)]
    assert min(rates) > 1.8, (errs, rates)
    rates3 = [np.log2(errs[3][i] / errs[3][i + 1]) for i in range(2)]
    assert min(rates3) > 1.8, (errs, rates3)
    assert all(e3 < e2 for e2, e3 in zip(errs[2], errs[3]))
