"""LUSGSMatrixVec and GMRES(k) with the LU-SGS preconditioner on the device.

* the matrix-vector product against the oracle (relative L2 <= 1e-12);
* GMRES against a dense solve of the same operator, assembled column by column from the
  oracle's matvec, with FP64 and FP32 preconditioner sweeps (the Krylov space and the
  operator stay FP64, so both reach the FP64 solution);
* one Krylov vector from zero is alpha M b with M the chain of single symmetric sweeps
  (1e-12): the preconditioner's skipped colour passes change nothing;
* the implicit dual-time step with gmresCode 1 / 2 reaches the same converged state as
  the LU-SGS path (<= 1e-10).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, random_state, rel_l2, small_mesh, vr_settings

pytestmark = pytest.mark.gpu


def _system(kind):
    import pyoracle as po
    import torch
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = small_mesh(kind, jitter=0.1)
    s = vr_settings(1)
    o = po.OracleVR(m, **s)
    g = VariationalReconstruction(m, s).construct()
    ev = EulerEvaluator(g, dict(riemannSolverType="Roe"))
    es = o.euler_settings(1.4, RS["Roe"])
    N = o.nCell
    u = random_state(N, 41)
    dt, lam, _, _ = o.eval_dt(es, u, 5.0)
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    du = g.to_device(u)
    dtd = torch.zeros(N, dtype=torch.float64, device=du.device)
    ev.EvaluateDt(dtd, du, None, 5.0)
    jdd = g.to_device(jd[:, None]).reshape(-1).contiguous()
    return o, g, ev, es, u, lam, jd, du, jdd


@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_lusgs_matvec_matches_oracle(require_gpu, kind):
    o, g, ev, es, u, lam, jd, du, jdd = _system(kind)
    x = np.random.default_rng(2).normal(size=(o.nCell, 5))
    import torch
    Ax = torch.empty_like(du)
    ev.LUSGSMatrixVec(1.0, 0.0, du, g.to_device(x), jdd, Ax)
    assert rel_l2(g.to_host(Ax), o.lusgs_matvec(es, 1.0, u, x, jd, lam)) <= 1e-12


@pytest.mark.parametrize("fp32", [False, True])
def test_gmres_solves_the_lusgs_system(require_gpu, fp32):
    import torch
    o, g, ev, es, u, lam, jd, du, jdd = _system("2d")
    N = o.nCell
    A = np.zeros((5 * N, 5 * N))
    for k in range(5 * N):
        e = np.zeros(5 * N)
        e[k] = 1.0
        A[:, k] = o.lusgs_matvec(es, 1.0, u, e.reshape(N, 5), jd, lam).ravel()
    b = np.random.default_rng(8).uniform(-1, 1, (N, 5))
    xTrue = np.linalg.solve(A, b.ravel()).reshape(N, 5)
    x = torch.zeros_like(du)
    res, resB = ev.GMRES(1.0, du, jdd, g.to_device(b), x, nGmresSpace=30, nGmresIter=6, sgsIter=1, fp32Precond=fp32)
    # Left preconditioning minimises |M (A x - b)|; an FP32 preconditioner is linear only to
    # FP32 rounding, so that residual (and the solution) floor near single precision.  In the
    # implicit loop the FP64 outer residual removes this floor (test below).
    tol = 1e-6 if fp32 else 1e-9
    assert res <= tol * resB
    xs = g.to_host(x)
    assert rel_l2(xs, xTrue) <= (1e-5 if fp32 else 1e-8)
    assert np.linalg.norm(A @ xs.ravel() - b.ravel()) <= (1e-5 if fp32 else 1e-8) * np.linalg.norm(b)


@pytest.mark.parametrize("sgsIter", [0, 2])
@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_gmres_one_krylov_vector_is_the_preconditioned_step(require_gpu, kind, sgsIter):
    """The preconditioner M of GMRES, (1 + sgsIter) symmetric Gauss-Seidel sweeps from zero (its
    colour passes that repeat the previous sweep's last one skipped), equals the chain of
    single UpdateSGS sweeps: one Krylov vector from x = 0 gives x = alpha M b with
    alpha = <M A M b, M b> / |M A M b|^2 (1e-12)."""
    import torch
    o, g, ev, es, u, lam, jd, du, jdd = _system(kind)
    b = g.to_device(np.random.default_rng(9).uniform(-1, 1, (o.nCell, 5)))

    def precond(r):
        a, c = torch.zeros_like(r), torch.zeros_like(r)
        for _ in range(sgsIter + 1):
            for fw in (True, False):
                ev.UpdateSGS(1.0, 0.0, r, du, a, c, jdd, fw, True)
                a, c = c, a
        return a

    Mb = precond(b)
    AMb = torch.empty_like(Mb)
    ev.LUSGSMatrixVec(1.0, 0.0, du, Mb, jdd, AMb)
    MAMb = precond(AMb)
    alpha = float((MAMb * Mb).sum()) / float((MAMb * MAMb).sum())
    x = torch.zeros_like(du)
    ev.GMRES(1.0, du, jdd, b, x, nGmresSpace=1, nGmresIter=1, sgsIter=sgsIter)
    assert rel_l2(g.to_host(x), g.to_host(alpha * Mb)) <= 1e-12

@pytest.mark.parametrize("gmresCode,fp32", [(1, False), (2, True)])
def test_implicit_step_with_gmres_converges_to_lusgs_state(require_gpu, gmresCode, fp32):
    from dndsr_amd import fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = mm.periodic_quad_grid(12, 10, 10.0, 10.0, jitter=0.1, seed=3)
    g = VariationalReconstruction(m, vr_settings(1, intOrder=3)).construct()
    ev = EulerEvaluator(g, dict(riemannSolverType="HLLEP"))
    ub = fields.iv_background_for_mach(0.5)
    u0 = g.cell_average(lambda p: fields.isentropic_vortex(p, 5.0, 1.4, ub), 5)
    import torch
    dtv = torch.zeros(g.nCell, dtype=torch.float64, device="cuda")
    dtE = ev.EvaluateDt(dtv, g.to_device(u0), None, 1.0, 1e100, False)
    out = {}
    for code in (0, gmresCode):
        du, dr = g.to_device(u0), g.BuildURec(5)
        n, res, rel, hist = ev.ImplicitEulerStep(du, dr, 8 * dtE, CFL=10.0, maxIter=600, tol=1e-12,
                                                 fp32Sweeps=fp32 and code != 0, gmresCode=code, nGmresSpace=10,
                                                 nGmresIter=2)
        assert n < 600 and rel[0] < 1e-12, (code, n, rel)
        out[code] = (g.to_host(du), n)
    assert rel_l2(out[gmresCode][0], out[0][0]) <= 1e-10
    assert out[gmresCode][1] <= out[0][1]
