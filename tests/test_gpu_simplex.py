"""Tri3 / Tet4 and mixed Tri3 + Quad4 meshes on the device vs the CPU oracle (FP64, relative L2 <= 1e-12 unless
noted): metrics and VR coefficients, the VR sweep, EvaluateDt, EvaluateRHS (inviscid and
viscous, several Riemann solvers), LU-SGS, SSP-RK3 and implicit steps on periodic and
bounded triangle and tetrahedron meshes (jittered lattices)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, colour_scan_order, random_rec, random_state, rel_l2, vr_settings

pytestmark = pytest.mark.gpu

TOL = 1e-12
GAMMA = 1.4
GAS = dict(Rgas=1.0, muGas=0.02, prGas=0.72, TRef=1.5, CSutherland=0.4)


def _cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])


def _mesh(kind):
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    vals = np.stack([_cons(1.0, [0.3, 0.1, 0.05], 1 / GAMMA), _cons(1.1, [0.4, 0.0, 0.05], 0.8),
                     _cons(0.9, [0.2, -0.1, 0.0], 0.75), _cons(1.05, [-0.1, 0.2, 0.1], 0.9)])
    if kind == "tri-periodic":
        return mm.tri_grid(7, 6, 3.5, 3.0, periodic=(True, True), jitter=0.15, seed=2), None, None
    if kind == "tet-periodic":
        return mm.tet_grid(4, 3, 3, 1.0, 0.8, 0.9, periodic=(True, True, True), jitter=0.12, seed=3), None, None
    if kind == "mixed-periodic":
        return mm.mixed_grid(8, 6, 3.5, 3.0, periodic=(True, True), jitter=0.15, seed=2), None, None
    if kind == "tri":
        m = mm.tri_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.12, seed=2)
        return m, [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM], vals
    if kind == "mixed":
        m = mm.mixed_grid(8, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.12, seed=2)
        return m, [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM], vals
    m = mm.tet_grid(4, 3, 3, 1.0, 0.8, 0.6, periodic=(False, True, False), zones=dict(xmin=0, xmax=1, zmin=2, zmax=3),
                    jitter=0.1, seed=3)
    return m, [bc.BC_IN, bc.BC_OUTP, bc.BC_WALL, bc.BC_WALL_INVIS], vals


CASES = [("tri-periodic", 1, 3), ("tri-periodic", 3, 5), ("tri", 2, 4), ("tet-periodic", 1, 3), ("tet-periodic", 2, 4),
         ("tet", 1, 3), ("tet", 2, 3), ("mixed-periodic", 1, 3), ("mixed-periodic", 3, 5), ("mixed", 2, 4)]


@pytest.fixture(scope="module", params=CASES, ids=[f"{k}-p{p}-i{i}" for k, p, i in CASES])
def pair(request, require_gpu):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    kind, order, intOrder = request.param
    print(f"\n[simplex parity] mesh={kind} p={order} intOrder={intOrder}", flush=True)
    m, types, vals = _mesh(kind)
    s = vr_settings(order, intOrder=intOrder)
    if types is None:
        o = po.OracleVR(m, **s)
        g = VariationalReconstruction(m, s).construct()
    else:
        zw = bc.euler_dir_weights(types, order)
        o = po.OracleVR(m, **s, zoneDirWeight=zw)
        o.set_bc(types, vals, GAMMA)
        g = VariationalReconstruction(m, dict(s, zoneDirWeight=zw)).construct()
    return m, o, g, types, vals


def _ev(g, types, vals, rs="Roe", viscous=False, muModel=0):
    import pyoracle as po
    from dndsr_amd.euler import EulerEvaluator
    es = po.OracleVR.euler_settings(GAMMA, RS[rs], 0, 1.0, 1.0, viscous=int(viscous), muModel=muModel, **GAS)
    st = dict(gamma=GAMMA, riemannSolverType=rs, viscous=viscous, idealGasProperty=dict(GAS, muModel=muModel))
    if types is not None:
        st.update(bcTypes=types, bcValues=vals)
    return es, EulerEvaluator(g, st)


def test_metrics_and_vr_coefficients(pair):
    m, o, g, _, _ = pair
    a, b = o.cell_metrics(), g.cell_metrics()
    for k in ("vol", "bary", "cent", "scale", "smooth"):
        assert rel_l2(b[k], a[k]) <= TOL, k
    # first moments vanish on simplices (the centre is the centroid): compare absolutely
    assert np.abs(b["moment"] - a["moment"]).max() <= 1e-12 * max(1.0, np.abs(a["moment"]).max())
    a, b = o.face_metrics(), g.face_metrics()
    for k in ("area", "meanNorm", "pts", "norms", "wJDet", "faceL", "faceWeight"):
        assert rel_l2(b[k], a[k]) <= TOL, k
    Ao, bo = o.vr_coeffs()
    Ag, bg = g.get_matrixAAInvB()
    assert rel_l2(Ag, Ao) <= 1e-10 and rel_l2(bg, bo) <= 1e-10


def test_vr_iter(pair):
    import torch
    m, o, g, types, vals = pair
    if types is not None:
        _ev(g, types, vals)          # sets the boundary table used by the VR boundary term
    Ao, bo = o.vr_coeffs()
    g.set_matrixAAInvB(Ao, bo)
    N, M = o.nCell, o.M
    u = random_state(N, seed=3)
    uRec = random_rec(N, M, 5, seed=4)
    du, dr = g.to_device(u), g.to_device(uRec)
    out = torch.zeros_like(dr)
    g.DoReconstructionIter(dr, out, du)
    assert rel_l2(g.to_host(out), o.vr_iter(uRec, u)) <= TOL


@pytest.mark.parametrize("rs,viscous", [("Roe", False), ("HLLEP", False), ("HLLC", False), ("Roe_M3", False),
                                        ("Roe", True)])
def test_eval_dt_and_rhs(pair, rs, viscous):
    m, o, g, types, vals = pair
    es, ev = _ev(g, types, vals, rs, viscous=viscous, muModel=1)
    N, M = o.nCell, o.M
    u = random_state(N, seed=11)
    uRec = random_rec(N, M, 5, seed=12)
    du, dr = g.to_device(u), g.to_device(uRec)
    dt = g.BuildUDof(1)[0]
    dtmin = ev.EvaluateDt(dt, du, None, CFL=0.8)
    dto, lamo, dlco, dmino = o.eval_dt(es, u, 0.8, 1e100, True)
    assert rel_l2(dt.cpu().numpy(), dto) <= TOL
    assert rel_l2(ev.lambdaFace.cpu().numpy().T, lamo) <= TOL
    assert abs(dtmin - dmino) <= TOL * dmino
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, du, dr, dr)
    assert rel_l2(g.to_host(rhs), o.eval_rhs(es, u, uRec, dlco)) <= TOL
    ev.EvaluateRHS(rhs, None, du, None, None)
    assert rel_l2(g.to_host(rhs), o.eval_rhs(es, u, None, dlco)) <= TOL


def test_lusgs(pair):
    m, o, g, types, vals = pair
    es, ev = _ev(g, types, vals)
    N = o.nCell
    u = random_state(N, seed=21)
    du = g.to_device(u)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, du, None, CFL=5.0)
    dto, lamo, _, _ = o.eval_dt(es, u, 5.0, 1e100, True)
    jd = g.BuildUDof(1)[0]
    ev.LUSGSMatrixInit(jd, None, dt, 0.3, 1.0, du, None)
    jdo = o.lusgs_init(dto, 0.3, 1.0, lamo)
    assert rel_l2(jd.cpu().numpy(), jdo) <= TOL
    rhs = np.random.default_rng(22).uniform(-1, 1, (N, 5))
    order = colour_scan_order(g.colors())
    dInc, dNew = g.BuildUDof(5), g.BuildUDof(5)
    ev.UpdateSGS(1.0, 0.0, g.to_device(rhs), du, dInc, dNew, jd, True, True, True)
    r1, _ = o.update_sgs(es, 1.0, rhs, u, np.zeros_like(u), np.zeros_like(u), jdo, lamo, True, True, True, order)
    assert rel_l2(g.to_host(dNew), r1) <= TOL


@pytest.mark.parametrize("kind,order,viscous", [("tri", 1, False), ("tet", 1, True), ("tet-periodic", 2, False),
                                                ("mixed", 2, True)])
def test_time_loops(require_gpu, kind, order, viscous):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    print(f"\n[simplex steps] {kind} p={order} viscous={viscous}", flush=True)
    m, types, vals = _mesh(kind)
    s = vr_settings(order, intOrder=3 if order == 1 else 4)
    if types is None:
        o = po.OracleVR(m, **s)
        g = VariationalReconstruction(m, s).construct()
    else:
        zw = bc.euler_dir_weights(types, order)
        o = po.OracleVR(m, **s, zoneDirWeight=zw)
        o.set_bc(types, vals, GAMMA)
        g = VariationalReconstruction(m, dict(s, zoneDirWeight=zw)).construct()
    es, ev = _ev(g, types, vals, "Roe", viscous=viscous)
    pts, wj = g.cell_quad()
    rho = 1.0 + 0.1 * np.sin(2 * pts[..., 0]) * np.cos(pts[..., 1])
    vel = np.stack([0.3 + 0.1 * np.sin(3 * pts[..., 1]), 0.05 * np.cos(2 * pts[..., 0]), 0.02 * np.sin(pts[..., 2])], -1)
    p = 1 / GAMMA * (1 + 0.05 * np.sin(3 * pts[..., 1]))
    U = np.concatenate([rho[..., None], rho[..., None] * vel,
                        (p / (GAMMA - 1) + 0.5 * rho * (vel**2).sum(-1))[..., None]], -1)
    u0 = np.einsum("ngv,ng->nv", U, wj) / wj.sum(1)[:, None]
    du, dr = g.to_device(u0), g.BuildURec(5)
    ev.SSPRK3(du, dr, 3, 0.4, 1)
    uo, ro = o.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 3, 0.4, 1)
    assert rel_l2(g.to_host(du), uo) <= 1e-11
    assert rel_l2(g.to_host(dr), ro) <= 1e-10
    dt = 5 * o.eval_dt(es, u0, 0.5, 1e100, False)[3]
    ordr = colour_scan_order(g.colors())
    du, dr = g.to_device(u0), g.BuildURec(5)
    n, res, rel, hist = ev.ImplicitEulerStep(du, dr, dt, CFL=10.0, maxIter=3, minIter=3, tol=0.0)
    ui, ri, ni, *_ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), dt, CFL=10.0, maxIter=3, minIter=3,
                                           tol=0.0, scanOrder=ordr)
    assert n == ni == 3
    assert rel_l2(g.to_host(du), ui) <= 1e-11
    assert rel_l2(g.to_host(dr), ri) <= 1e-10
