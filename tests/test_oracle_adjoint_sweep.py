"""CPU pins of the oracle's transposed LU-SGS sweep (the adjoint preconditioner): the dense
LU-SGS matrix P is assembled column by column from the oracle's LUSGSMatrixVec
(EulerEvaluator.hxx:175-275), and the transposed sweep must be a Jacobi or Gauss-Seidel
sweep of P^T: y_new = D^-1 (b - (P^T - D) y) with, for Gauss-Seidel in scan order, the
entries of already-visited cells taken from y_new."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import bc
from dndsr_amd import mesh as mm
from helpers import RS, random_state


def _case(kind):
    if kind == "periodic":
        m = mm.periodic_quad_grid(5, 4, 5.0, 4.0, jitter=0.15, seed=2)
        o = po.OracleVR(m, maxOrder=1, intOrder=3)
    else:
        types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
        m = mm.box_quad_grid(5, 4, 5.0, 4.0, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.15, seed=2)
        o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
        o.set_bc(types, np.tile(random_state(1, seed=9)[0], (3, 1)), 1.4)
    es = o.euler_settings(1.4, RS["Roe"])
    u = random_state(o.nCell, seed=4)
    dTau, lam, _, _ = o.eval_dt(es, u, 2.0, 1e100, True)
    jd = o.lusgs_init(dTau, 0.7, 1.0, lam)
    return o, es, u, jd, lam


@pytest.mark.parametrize("kind", ["periodic", "bounded"])
@pytest.mark.parametrize("gsUpdate,forward", [(0, True), (1, True), (1, False)])
def test_transposed_sweep_is_a_sweep_of_the_transpose(kind, gsUpdate, forward):
    o, es, u, jd, lam = _case(kind)
    N = o.nCell
    n = 5 * N
    P = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        P[:, c] = o.lusgs_matvec(es, 1.0, u, e.reshape(N, 5), jd, lam).ravel()
    PT = P.T
    D = np.repeat(jd, 5)
    rng = np.random.default_rng(3)
    b, y0 = rng.normal(size=n), rng.normal(size=n)
    out, s = o.update_sgs_transpose(es, 1.0, b.reshape(N, 5), u, y0.reshape(N, 5), np.zeros((N, 5)), jd, lam,
                                    forward, gsUpdate, False)
    ref = np.zeros(n)
    cells = range(N) if forward else range(N - 1, -1, -1)
    for i in cells:
        rows = slice(5 * i, 5 * i + 5)
        y = y0.copy()
        if gsUpdate:  # already-visited cells take the new values
            done = np.arange(N) < i if forward else np.arange(N) > i
            mask = np.repeat(done, 5)
            y[mask] = ref[mask]
        off = PT[rows] @ y - D[rows] * y[rows]
        ref[rows] = (b[rows] - off) / D[rows]
    assert np.abs(out.ravel() - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_allclose(s, np.abs(ref.reshape(N, 5)).sum(0), rtol=1e-12)


def _block_case(kind):
    if kind == "periodic":
        m = mm.periodic_quad_grid(5, 4, 5.0, 4.0, jitter=0.15, seed=2)
        o = po.OracleVR(m, maxOrder=1, intOrder=3)
    else:
        types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
        m = mm.box_quad_grid(5, 4, 5.0, 4.0, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.15, seed=2)
        o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
        o.set_bc(types, np.tile(random_state(1, seed=9)[0], (3, 1)), 1.4)
    es = o.euler_settings(1.4, RS["Roe"], viscous=1, useRoeJacobian=1, Rgas=1.0, muGas=0.05, prGas=0.72,
                          muModel=0)
    u = random_state(o.nCell, seed=4)
    rec = 0.05 * np.random.default_rng(5).normal(size=(o.nCell, 5, o.M))
    dTau, lam, dlc, _ = o.eval_dt(es, u, 2.0, 1e100, True)
    _, lr = o.eval_rhs(es, u, rec, dlc, roeLambdas=True)
    jd, ji = o.lusgs_init_block(es, dTau, 0.7, 1.0, u, lam, lr)
    return o, es, u, jd, ji, lam, lr


@pytest.mark.parametrize("kind", ["periodic", "bounded"])
@pytest.mark.parametrize("gsUpdate,forward", [(0, True), (1, True), (1, False)])
def test_transposed_block_sweep_is_a_sweep_of_the_transpose(kind, gsUpdate, forward):
    """The block matrix M = JDiag + O is read off the forward block sweep (a Jacobi sweep
    with a zero right-hand side returns -JDiag^-1 O x), then the transposed block sweep must be
    a Jacobi or Gauss-Seidel sweep of M^T with JDiag^T on the diagonal."""
    o, es, u, jd, ji, lam, lr = _block_case(kind)
    N = o.nCell
    n = 5 * N
    Off = np.zeros((n, n))
    zero = np.zeros((N, 5))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        out, _ = o.update_sgs_block(es, 1.0, zero, u, e.reshape(N, 5), zero, ji, lam, lr, True, 0, False)
        Off[:, c] = -np.einsum("nij,nj->ni", jd, out).ravel()
    assert np.abs(Off).max() > 0
    Mt = Off.T.copy()
    for i in range(N):
        Mt[5 * i:5 * i + 5, 5 * i:5 * i + 5] = jd[i].T
    rng = np.random.default_rng(3)
    b, y0 = rng.normal(size=n), rng.normal(size=n)
    out, s = o.update_sgs_block_transpose(es, 1.0, b.reshape(N, 5), u, y0.reshape(N, 5), zero, ji, lam, lr,
                                          forward, gsUpdate, False)
    ref = np.zeros(n)
    cells = range(N) if forward else range(N - 1, -1, -1)
    for i in cells:
        rows = slice(5 * i, 5 * i + 5)
        y = y0.copy()
        if gsUpdate:
            done = np.arange(N) < i if forward else np.arange(N) > i
            mask = np.repeat(done, 5)
            y[mask] = ref[mask]
        y[rows] = 0
        ref[rows] = np.linalg.solve(jd[i].T, b[rows] - Mt[rows] @ y)
    assert np.abs(out.ravel() - ref).max() <= 1e-11 * np.abs(ref).max()
    np.testing.assert_allclose(s, np.abs(ref.reshape(N, 5)).sum(0), rtol=1e-11)
