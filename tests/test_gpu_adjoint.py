"""Discrete adjoint on the device, checked against finite differences.

* EvaluateRHS VJP: <lambda, (R(x + e v) - R(x - e v)) / 2e> against <VJP(lambda), v> for
  random directions in (u, uRec), every Riemann solver family, 2-D and 3-D,
  relative difference <= 1e-6 (the north-star bound for adjoint gradients).
* VR sweep VJP: the sweep is linear in (uRec, u), so the dot-product identity
  <ybar, S(x, u)> = <S^T ybar, (x, u)> holds to rounding (1e-12).
* ResidualFunction (nRec sweeps + RHS, reconstruction recomputed in backward):
  torch.autograd.gradcheck on a tiny mesh.
* Functional gradient: dJ/dtheta of J = sum_i V_i w_i . rhs_i(u(theta)) from the
  device adjoint against the central FD (step 1e-6) of the same functional evaluated
  with the CPU oracle (the reference-path forward operators).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, random_rec, random_state, small_mesh, vr_settings

pytestmark = pytest.mark.gpu


def _setup(kind, order, rs):
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = small_mesh(kind, jitter=0.1)
    s = vr_settings(order)
    g = VariationalReconstruction(m, s).construct()
    batch = rs == "Roe_M1b"
    ev = EulerEvaluator(g, dict(riemannSolverType="Roe_M1" if batch else rs, rsMeanValueEig=1 if batch else 0))
    return m, s, g, ev


def _dot(a, b):
    return float((a.double() * b.double()).sum())


@pytest.mark.parametrize("kind,order", [("2d", 1), ("2d", 2), ("2d", 3), ("3d", 1), ("3d", 2), ("3d", 3)])
@pytest.mark.parametrize("rs", ["Roe", "HLLEP", "HLLC", "Roe_M1b"])
def test_rhs_vjp_matches_central_fd(require_gpu, kind, order, rs):
    import torch
    from dndsr_amd import adjoint
    print(f"\n[adjoint] rhs vjp {kind} p{order} {rs}", flush=True)
    m, s, g, ev = _setup(kind, order, rs)
    N, M = g.nCell, g.M
    u = g.to_device(random_state(N, 5))
    rec = g.to_device(random_rec(N, M, 5, seed=6, amp=2e-2))
    dt = torch.zeros(N, dtype=torch.float64, device=u.device)
    ev.EvaluateDt(dt, u, None, 1.0)
    dlc = ev.deltaLambdaCell.clone()
    gen = torch.Generator(device="cpu").manual_seed(3)
    lam = torch.randn(u.shape, generator=gen, dtype=torch.float64).to(u.device)
    vu = torch.randn(u.shape, generator=gen, dtype=torch.float64).to(u.device) * 1e-1
    vr = torch.randn(rec.shape, generator=gen, dtype=torch.float64).to(u.device) * 1e-2
    uBar, recBar = adjoint.rhs_vjp(ev, u, rec, dlc, lam)
    lhs = _dot(uBar, vu) + _dot(recBar, vr)
    eps = 1e-6
    rp = adjoint._eval_rhs(ev, u + eps * vu, rec + eps * vr, dlc)
    rm = adjoint._eval_rhs(ev, u - eps * vu, rec - eps * vr, dlc)
    fd = _dot(lam, (rp - rm) / (2 * eps))
    assert abs(lhs - fd) <= 1e-6 * max(abs(fd), 1e-30), (lhs, fd)


@pytest.mark.parametrize("kind,order", [("2d", 1), ("2d", 3), ("3d", 2)])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_vr_vjp_dot_product(require_gpu, kind, order, mode):
    import torch
    from dndsr_amd import adjoint
    from dndsr_amd.cfv import VariationalReconstruction
    m = small_mesh(kind, jitter=0.1)
    g = VariationalReconstruction(m, vr_settings(order)).construct()
    N, M = g.nCell, g.M
    gen = torch.Generator(device="cpu").manual_seed(11)
    x = torch.randn((5, M, N), generator=gen, dtype=torch.float64).cuda()
    u = torch.randn((5, N), generator=gen, dtype=torch.float64).cuda()
    yb = torch.randn((5, M, N), generator=gen, dtype=torch.float64).cuda()
    y = torch.empty_like(x)
    if mode == 3:
        g.DoReconstructionIterDiff(None, x, y)
    else:
        g.DoReconstructionIter(x, y, u, putIntoNew=True, recordInc=mode >= 1, uRecIsZero=mode == 2, relax=0.8)
    xb, ub = adjoint.vr_vjp(g, yb, mode, 0.8)
    lhs = _dot(yb, y)
    rhs = _dot(xb, x) + (0.0 if ub is None else _dot(ub, u))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + _dot(yb.abs(), y.abs()))


def test_residual_function_gradcheck(require_gpu):
    import torch
    from dndsr_amd import adjoint
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = mm.periodic_quad_grid(4, 3, 4.0, 3.0, jitter=0.1, seed=2)
    g = VariationalReconstruction(m, vr_settings(1)).construct()
    ev = EulerEvaluator(g, dict(riemannSolverType="HLLEP"))
    N, M = g.nCell, g.M
    u = g.to_device(random_state(N, 8))
    dt = torch.zeros(N, dtype=torch.float64, device=u.device)
    ev.EvaluateDt(dt, u, None, 1.0)
    dlc = ev.deltaLambdaCell.clone()
    rec0 = g.to_device(random_rec(N, M, 5, seed=9, amp=1e-2))
    u.requires_grad_(True)
    rec0.requires_grad_(True)
    f = lambda a, b: adjoint.ResidualFunction.apply(a, b, ev, 2, dlc)  # noqa: E731
    assert torch.autograd.gradcheck(f, (u, rec0), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("rs", ["Roe", "HLLEP"])
def test_functional_gradient_matches_oracle_fd(require_gpu, rs):
    """J(theta) = sum_i V_i w_i . frhs_i(u(theta)), u(theta) = cell means of an isentropic
    vortex of strength theta, frhs = 2 Jacobi VR sweeps from zero + EvaluateRHS, with
    deltaLambdaCell frozen at theta0.  Device adjoint vs oracle central FD (step 1e-6)."""
    import torch
    import pyoracle as po
    from dndsr_amd import adjoint, fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = mm.periodic_quad_grid(10, 9, 10.0, 10.0, jitter=0.1, seed=4)
    s = vr_settings(1, intOrder=3)
    g = VariationalReconstruction(m, s).construct()
    o = po.OracleVR(m, **s)
    ev = EulerEvaluator(g, dict(riemannSolverType=rs))
    es = o.euler_settings(1.4, RS[rs])
    ub = fields.iv_background_for_mach(0.5)
    uOf = lambda th: g.cell_average(lambda p: fields.isentropic_vortex(p, th, 1.4, ub), 5)  # noqa: E731
    th0, h = 5.0, 1e-6
    u0 = uOf(th0)
    N, M = g.nCell, g.M
    w = np.random.default_rng(12).uniform(-1, 1, (N, 5))
    vol = o.cell_metrics()["vol"]
    _, _, dlc, _ = o.eval_dt(es, u0, 1.0, 1e100, True)

    def J_oracle(th):
        uu = uOf(th)
        rec = np.zeros((N, 5, M))
        for _ in range(2):
            rec = o.vr_iter(rec, uu)
        return float((vol[:, None] * w * o.eval_rhs(es, uu, rec, dlc)).sum())

    fd = (J_oracle(th0 + h) - J_oracle(th0 - h)) / (2 * h)
    du = g.to_device(u0).requires_grad_(True)
    rec0 = g.BuildURec(5)
    lamW = g.to_device(vol[:, None] * w)
    rhs = adjoint.residual(ev, du, rec0, 2, g.to_device(dlc[:, None]).reshape(-1).contiguous())
    (rhs * lamW).sum().backward()
    dudth = (uOf(th0 + h) - uOf(th0 - h)) / (2 * h)
    grad = float((g.to_host(du.grad) * dudth).sum())
    assert abs(grad - fd) <= 1e-6 * abs(fd), (grad, fd)


def test_library_calls_are_ordered_with_torch_ops(require_gpu):
    """Library kernels run on torch's current stream: a tensor written by torch and read by
    the library (and the reverse) needs no explicit synchronisation."""
    import torch
    from dndsr_amd import adjoint
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m = mm.periodic_hex_grid(24, 24, 24, 1.0, 1.0, 1.0)
    g = VariationalReconstruction(m, vr_settings(2, intOrder=5)).construct()
    ev = EulerEvaluator(g, dict(riemannSolverType="HLLEP"))
    u0 = g.to_device(random_state(g.nCell, 21))
    rec = g.to_device(random_rec(g.nCell, g.M, 5, seed=22))
    dlc = torch.zeros(g.nCell, dtype=torch.float64, device=u0.device)
    ref = adjoint._eval_rhs(ev, u0, rec, dlc)
    torch.cuda.synchronize()
    for k in range(5):
        u = u0 * 1.0 + 0.0  # fresh torch-written tensor every round, no synchronisation
        r = adjoint._eval_rhs(ev, u, rec, dlc)
        assert torch.equal(r, ref), k
