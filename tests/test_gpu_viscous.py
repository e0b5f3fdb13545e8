"""Laminar viscous terms on the device vs the CPU oracle (FP64): EvaluateDt with the
viscous spectral radius, EvaluateRHS with the viscous face fluxes (its viscous part,
viscous minus inviscid residual, to 1e-10 of that part; the full residual to 1e-12), for
the three viscosity laws, periodic and bounded meshes (no-slip adiabatic walls, symmetry,
far field, outflow, inflow, back-pressure outflow), reconstruction orders 1-3, and the
device time loops (SSP-RK3, implicit dual-time with LU-SGS) on a viscous run.  The viscous
adjoint is checked in test_gpu_adjoint_bc.py."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, random_rec, random_state, rel_l2, small_mesh, vr_settings

pytestmark = pytest.mark.gpu

TOL = 1e-12
GAMMA = 1.4
GAS = dict(Rgas=1.0, muGas=0.05, prGas=0.72, TRef=1.5, CSutherland=0.4)


def _cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])


def _mesh(kind):
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    if kind in ("2d-periodic", "3d-periodic"):
        return small_mesh(kind[:2]), None, None
    if kind == "2d":
        m = mm.box_quad_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.12, seed=2)
        types = [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM]
    else:
        m = mm.box_hex_grid(5, 4, 3, 1.0, 0.8, 0.6, periodic=(False, True, False),
                            zones=dict(xmin=0, xmax=1, zmin=2, zmax=3), jitter=0.1, seed=3)
        types = [bc.BC_IN, bc.BC_OUTP, bc.BC_WALL, bc.BC_WALL_INVIS]
    vals = np.stack([_cons(1.0, [0.3, 0.1, 0.0], 1 / GAMMA), _cons(1.1, [0.4, 0.0, 0.05], 0.8),
                     _cons(0.9, [0.2, -0.1, 0.0], 0.75), _cons(1.05, [-0.1, 0.2, 0.1], 0.9)])
    return m, types, vals


CASES = [("2d-periodic", 1), ("2d-periodic", 3), ("3d-periodic", 2), ("2d", 1), ("2d", 2), ("3d", 1), ("3d", 2),
         ("3d", 3)]


@pytest.fixture(scope="module", params=CASES, ids=[f"{k}-p{p}" for k, p in CASES])
def pair(request, require_gpu):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    kind, order = request.param
    print(f"\n[viscous parity] mesh={kind} p={order}", flush=True)
    m, types, vals = _mesh(kind)
    s = vr_settings(order)
    if types is None:
        o = po.OracleVR(m, **s)
        g = VariationalReconstruction(m, s).construct()
    else:
        zw = bc.euler_dir_weights(types, order)
        o = po.OracleVR(m, **s, zoneDirWeight=zw)
        o.set_bc(types, vals, GAMMA)
        g = VariationalReconstruction(m, dict(s, zoneDirWeight=zw)).construct()
    return m, o, g, types, vals


def _ev(g, types, vals, rs="Roe", viscous=True, muModel=0):
    import pyoracle as po
    from dndsr_amd.euler import EulerEvaluator
    es = po.OracleVR.euler_settings(GAMMA, RS[rs], 0, 1.0, 1.0, viscous=int(viscous), muModel=muModel, **GAS)
    st = dict(gamma=GAMMA, riemannSolverType=rs, viscous=viscous, idealGasProperty=dict(GAS, muModel=muModel))
    if types is not None:
        st.update(bcTypes=types, bcValues=vals)
    return es, EulerEvaluator(g, st)


@pytest.mark.parametrize("muModel", [0, 1, 2])
def test_eval_dt_viscous(pair, muModel):
    m, o, g, types, vals = pair
    es, ev = _ev(g, types, vals, muModel=muModel)
    u = random_state(o.nCell, seed=5)
    du = g.to_device(u)
    dt = g.BuildUDof(1)[0]
    dtmin = ev.EvaluateDt(dt, du, None, CFL=0.7, UseLocaldt=True)
    dto, lamo, dlco, dmino = o.eval_dt(es, u, 0.7, 1e100, True)
    assert rel_l2(dt.cpu().numpy(), dto) <= TOL
    assert rel_l2(ev.lambdaFace.cpu().numpy().T, lamo) <= TOL
    assert abs(dtmin - dmino) <= TOL * dmino
    assert np.abs(lamo[:, 2]).max() > 0.05 * np.abs(lamo[:, 0]).max()   # the viscous radius matters


@pytest.mark.parametrize("rs,muModel", [("Roe", 0), ("Roe", 1), ("HLLEP", 2), ("HLLC", 0)])
def test_eval_rhs_viscous(pair, rs, muModel):
    m, o, g, types, vals = pair
    es, ev = _ev(g, types, vals, rs, muModel=muModel)
    es_i, ev_i = _ev(g, types, vals, rs, viscous=False)
    N, M = o.nCell, o.M
    u = random_state(N, seed=11)
    uRec = random_rec(N, M, 5, seed=12)
    du, dr = g.to_device(u), g.to_device(uRec)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, du, None, CFL=1.0)
    ev_i.EvaluateDt(dt, du, None, CFL=1.0)
    _, _, dlco, _ = o.eval_dt(es, u, 1.0, 1e100, True)
    rhs, rhs_i = g.BuildUDof(5), g.BuildUDof(5)
    for rec, drec in ((uRec, dr), (None, None)):
        ev.EvaluateRHS(rhs, None, du, drec, drec)
        ev_i.EvaluateRHS(rhs_i, None, du, drec, drec)
        ro, ro_i = o.eval_rhs(es, u, rec, dlco), o.eval_rhs(es_i, u, rec, dlco)
        assert rel_l2(g.to_host(rhs), ro) <= TOL
        vis_o = ro - ro_i
        assert np.abs(vis_o).max() > 1e-3 * np.abs(ro).max()
        assert rel_l2(g.to_host(rhs) - g.to_host(rhs_i), vis_o) <= 1e-10


def test_viscous_time_loops(require_gpu):
    """SSP-RK3 and implicit dual-time steps of a viscous run on a bounded 2-D mesh (wall,
    symmetry, far field, outflow) and a periodic 3-D mesh."""
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    from helpers import colour_scan_order
    for kind, order in (("2d", 1), ("3d-periodic", 2)):
        print(f"\n[viscous steps] {kind} p={order}", flush=True)
        m, types, vals = _mesh(kind)
        s = vr_settings(order)
        if types is None:
            o = po.OracleVR(m, **s)
            g = VariationalReconstruction(m, s).construct()
        else:
            zw = bc.euler_dir_weights(types, order)
            o = po.OracleVR(m, **s, zoneDirWeight=zw)
            o.set_bc(types, vals, GAMMA)
            g = VariationalReconstruction(m, dict(s, zoneDirWeight=zw)).construct()
        es, ev = _ev(g, types, vals, "Roe", muModel=1)
        pts, wj = g.cell_quad()
        rho = 1.0 + 0.1 * np.sin(2 * pts[..., 0]) * np.cos(pts[..., 1])
        vel = np.stack([0.3 + 0.1 * np.sin(3 * pts[..., 1]), 0.05 * np.cos(2 * pts[..., 0]), 0 * rho], -1)
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
        ui, ri, ni, *_ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), dt, CFL=10.0, maxIter=3,
                                               minIter=3, tol=0.0, scanOrder=ordr)
        assert n == ni == 3
        assert rel_l2(g.to_host(du), ui) <= 1e-11
        assert rel_l2(g.to_host(dr), ri) <= 1e-10
