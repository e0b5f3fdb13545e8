"""Second-order substitute of the VR sweep on the device (maxOrder 1, subs2ndOrder 1 = Gauss
rule with both face interpolations, 2 = least squares) vs the CPU oracle (FP64; relative L2
<= 1e-12 for the Gauss rule, 1e-11 for least squares, which the device solves by normal
equations and the oracle by Householder QR): bounded meshes with Euler ghost states,
periodic faces (face centre and barycentre seen from the front cell), nVars 1 on a
periodic mesh, triangles, mixed cells and tetrahedra; every dndsr_vr_iter mode dispatches
to it, the Jacobian action is zero, the transpose against central differences; SSP-RK3 and
implicit steps with the substitute."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import colour_scan_order, random_state, rel_l2, small_mesh

pytestmark = pytest.mark.gpu

GAMMA = 1.4
TOL = {1: 1e-12, 2: 1e-11}


def _pair(m, method, scheme, types=None, vals=None):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    s = dict(maxOrder=1, intOrder=3, subs2ndOrder=method, subs2ndOrderGGScheme=scheme)
    if types is not None:
        s["zoneDirWeight"] = bc.euler_dir_weights(types, 1)
    o = po.OracleVR(m, **s)
    g = VariationalReconstruction(m, s).construct()
    ev = None
    if types is not None:
        o.set_bc(types, vals, GAMMA)
        ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=vals))
    return o, g, ev


def _mesh(kind):
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from test_gpu_bc import _case
    if kind in ("2d", "3d"):
        return _case(kind)
    types = [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM]
    vals = np.tile([1.0, 0.3, 0.05, 0.0, 1 / 0.56 + 0.5 * 0.0925], (4, 1))
    z2 = dict(xmin=0, xmax=1, ymin=2, ymax=3)
    if kind == "tri":
        return mm.tri_grid(7, 6, 3.5, 3.0, zones=z2, jitter=0.1, seed=2), types, vals
    if kind == "mixed":
        return mm.mixed_grid(7, 6, 3.5, 3.0, zones=z2, jitter=0.1, seed=2), types, vals
    return mm.tet_grid(4, 3, 3, 1.0, 0.8, 0.6, zones=dict(z2, zmin=2, zmax=3), jitter=0.1, seed=3), types, vals


CASES = [("2d", 1, 1), ("2d", 1, 0), ("2d", 2, 1), ("3d", 1, 1), ("3d", 2, 1), ("tri", 1, 1), ("mixed", 2, 1),
         ("tet", 1, 1), ("tet", 2, 1)]


@pytest.mark.parametrize("kind,method,scheme", CASES, ids=[f"{k}-m{m}-s{s}" for k, m, s in CASES])
def test_substitute_matches_oracle(require_gpu, kind, method, scheme):
    import torch
    print(f"\n[2nd] {kind} subs2ndOrder={method} scheme={scheme}", flush=True)
    m, types, vals = _mesh(kind)
    o, g, ev = _pair(m, method, scheme, types, vals)
    N = o.nCell
    u = random_state(N, seed=21)
    ref = o.vr_2nd(u)
    assert np.abs(ref).max() > 1e-3
    du = g.to_device(u)
    rec = g.to_device(np.random.default_rng(3).uniform(-1, 1, ref.shape))
    out = torch.zeros_like(rec)
    for kw in (dict(), dict(recordInc=True), dict(recordInc=True, uRecIsZero=True)):
        out.zero_()
        g.DoReconstructionIter(rec, out, du, **kw)
        assert rel_l2(g.to_host(out), ref) <= TOL[method], kw
    g.DoReconstructionSOR(rec, du)
    assert rel_l2(g.to_host(rec), ref) <= TOL[method]
    out.fill_(1.0)
    g.DoReconstructionIterDiff(rec, torch.ones_like(rec), out, du)
    assert float(out.abs().max()) == 0.0


@pytest.mark.parametrize("method", [1, 2])
def test_substitute_periodic_one_variable(require_gpu, method):
    """nVars 1 on a jittered periodic mesh: every cell sees periodic faces from both sides."""
    m = small_mesh("2d")
    o, g, _ = _pair(m, method, 1)
    u = np.random.default_rng(4).uniform(-1, 1, (o.nCell, 1))
    dr = g.BuildURec(1)
    out = g.BuildURec(1)
    g.DoReconstructionIter(dr, out, g.to_device(u))
    assert rel_l2(g.to_host(out), o.vr_2nd(u)) <= TOL[method]


@pytest.mark.parametrize("kind,method,scheme", CASES, ids=[f"{k}-m{m}-s{s}" for k, m, s in CASES])
def test_substitute_transpose_matches_central_fd(require_gpu, kind, method, scheme):
    """The transposed substitute (dndsr_vr_iter_vjp with subs2ndOrder): nothing flows back to
    uRec, and <ybar, dG/du v> from central differences equals <G^T ybar, v> (relative 1e-6, the
    bound of the other transposes; the substitute is linear but for the boundary ghosts)."""
    import torch
    from dndsr_amd.adjoint import _vr_sweep, vr_vjp
    m, types, vals = _mesh(kind)
    o, g, ev = _pair(m, method, scheme, types, vals)
    N = o.nCell
    du = g.to_device(random_state(N, seed=2))
    gen = torch.Generator(device="cpu").manual_seed(5)
    yb = torch.randn((5, g.M, N), generator=gen, dtype=torch.float64).to(du.device)
    v = torch.randn(du.shape, generator=gen, dtype=torch.float64).to(du.device) * 0.05
    rec = g.BuildURec(5)
    xb, ub = vr_vjp(g, yb, 0, 1.0, rec, du)
    assert float(xb.abs().max()) == 0.0
    eps = 1e-6
    fd = float((yb * (_vr_sweep(g, rec, du + eps * v, 1.0) - _vr_sweep(g, rec, du - eps * v, 1.0))).sum()) / (2 * eps)
    lhs = float((ub * v).sum())
    assert abs(lhs - fd) <= 1e-6 * abs(fd), (lhs, fd)


@pytest.mark.parametrize("method", [1, 2])
def test_substitute_transpose_periodic_dot_product(require_gpu, method):
    """On a periodic mesh with nVars 1 the substitute is linear in u: <ybar, G v> equals
    <G^T ybar, v> to rounding (relative 1e-12), and the transpose is linear in ybar."""
    import torch
    from dndsr_amd.adjoint import _vr_sweep, vr_vjp
    m = small_mesh("2d")
    o, g, _ = _pair(m, method, 1)
    N = o.nCell
    gen = torch.Generator(device="cpu").manual_seed(6)
    rec = g.BuildURec(1)
    v = torch.randn((1, N), generator=gen, dtype=torch.float64).to(rec.device)
    y1 = torch.randn((1, g.M, N), generator=gen, dtype=torch.float64).to(rec.device)
    y2 = torch.randn((1, g.M, N), generator=gen, dtype=torch.float64).to(rec.device)
    Gv = _vr_sweep(g, rec, v, 1.0)
    _, b1 = vr_vjp(g, y1, 0, 1.0, rec, v)
    _, b2 = vr_vjp(g, y2, 0, 1.0, rec, v)
    _, b12 = vr_vjp(g, 2.0 * y1 - y2, 0, 1.0, rec, v)
    lhs, rhs = float((y1 * Gv).sum()), float((b1 * v).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    assert float((b12 - (2.0 * b1 - b2)).abs().max()) <= 1e-12 * float(b1.abs().max())


@pytest.mark.parametrize("method", [1, 2])
def test_solver_loops_with_substitute(require_gpu, method):
    import pyoracle as po
    from test_gpu_bc import _flow
    print(f"\n[2nd steps] subs2ndOrder={method}", flush=True)
    m, types, vals = _mesh("2d")
    o, g, ev = _pair(m, method, 1, types, vals)
    es = po.OracleVR.euler_settings(GAMMA, 1, 0, 1.0, 1.0)
    u0 = _flow(o, g, "2d")
    du, dr = g.to_device(u0), g.BuildURec(5)
    ev.SSPRK3(du, dr, 3, 0.4, 1)
    uo, ro = o.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 3, 0.4, 1)
    assert rel_l2(g.to_host(du), uo) <= 1e-11
    assert rel_l2(g.to_host(dr), ro) <= 1e-10
    dt = 5 * o.eval_dt(es, u0, 0.5, 1e100, False)[3]
    du, dr = g.to_device(u0), g.BuildURec(5)
    n, *_ = ev.ImplicitEulerStep(du, dr, dt, CFL=10.0, maxIter=3, minIter=3, tol=0.0)
    ui, ri, ni, *_ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), dt, CFL=10.0, maxIter=3,
                                           minIter=3, tol=0.0, scanOrder=colour_scan_order(g.colors()))
    assert n == ni == 3
    assert rel_l2(g.to_host(du), ui) <= 1e-11
    assert rel_l2(g.to_host(dr), ri) <= 1e-10
