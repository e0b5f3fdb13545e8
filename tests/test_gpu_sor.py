"""In-place SOR reconstruction sweep on the device (dndsr_vr_iter mode 4, SORInstead) vs
the CPU oracle's serial SOR in the device's colour order (FP64, relative L2 <= 1e-12):
periodic, bounded (Euler ghost states; boundary values given as data) and simplex/mixed
meshes, relaxed and unrelaxed, several sweeps; the reference's putIntoNew swap; the
device solver loops (SSP-RK3, implicit dual time) with SORInstead on colour-sorted
meshes, where the oracle's natural-order loops are the same sweep."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import colour_scan_order, random_rec, random_state, rel_l2, small_mesh, vr_settings

pytestmark = pytest.mark.gpu

TOL = 1e-12
GAMMA = 1.4


def _bounded(kind):
    from test_gpu_bc import _case
    return _case(kind)


def _pair(m, order, types=None, vals=None, relax=1.0, sor=False):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    s = vr_settings(order)
    s["jacobiRelax"] = relax
    if types is not None:
        s["zoneDirWeight"] = bc.euler_dir_weights(types, order)
    o = po.OracleVR(m, **s, SORInstead=sor)
    g = VariationalReconstruction(m, dict(s, SORInstead=sor)).construct()
    ev = None
    if types is not None:
        o.set_bc(types, vals, GAMMA)
        ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=vals))
    return o, g, ev


CASES = [("2d-periodic", 1, 1.0), ("2d-periodic", 3, 0.8), ("3d-periodic", 2, 1.0), ("3d-periodic", 3, 1.0),
         ("2d", 2, 1.0), ("2d", 1, 0.7), ("3d", 1, 1.0), ("3d", 2, 0.9)]


@pytest.mark.parametrize("kind,order,relax", CASES, ids=[f"{k}-p{p}-w{w}" for k, p, w in CASES])
def test_vr_sor_matches_oracle(require_gpu, kind, order, relax):
    print(f"\n[sor] {kind} p={order} relax={relax}", flush=True)
    if kind.endswith("periodic"):
        o, g, _ = _pair(small_mesh(kind[:2]), order, relax=relax)
    else:
        m, types, vals = _bounded(kind)
        o, g, ev = _pair(m, order, types, vals, relax)   # ev holds the boundary table
    Ao, bo = o.vr_coeffs()
    g.set_matrixAAInvB(Ao, bo)
    N, M = o.nCell, o.M
    u, rec = random_state(N, seed=3), random_rec(N, M, 5, seed=4)
    ordr = colour_scan_order(g.colors())
    du, dr = g.to_device(u), g.to_device(rec)
    ref = rec
    for sweep in range(3):
        g.DoReconstructionSOR(dr, du)
        ref = o.vr_sor(ref, u, ordr)
        assert rel_l2(g.to_host(dr), ref) <= TOL, sweep
    assert rel_l2(ref, o.vr_iter(rec, u)) > 1e-6        # not the Jacobi sweep


def test_vr_sor_point_values_and_put_into_new(require_gpu):
    """nVars = 1 with boundary values given as data; the putIntoNew swap of the arrays."""
    import torch
    m, _, _ = _bounded("2d")
    o, g, _ = _pair(m, 2, relax=0.9, sor=True)
    Ao, bo = o.vr_coeffs()
    g.set_matrixAAInvB(Ao, bo)
    N, M, F, G = o.nCell, o.M, o.nFace, o.nGFace
    rng = np.random.default_rng(8)
    val = rng.uniform(-1, 1, (F, G, 1))
    o.set_vr_dirichlet(val)
    g.set_vr_dirichlet(torch.from_numpy(np.ascontiguousarray(val.transpose(2, 1, 0))).to(g.ctx.torch_device))
    u, rec, other = rng.uniform(-1, 1, (N, 1)), random_rec(N, M, 1, seed=10), random_rec(N, M, 1, seed=11)
    ordr = colour_scan_order(g.colors())
    dr, dn, du = g.to_device(rec), g.to_device(other), g.to_device(u)
    g.DoReconstructionIter(dr, dn, du, putIntoNew=True)
    assert rel_l2(g.to_host(dn), o.vr_sor(rec, u, ordr)) <= TOL
    assert np.array_equal(g.to_host(dr), other)
    g.DoReconstructionIter(dr, dn, du, putIntoNew=False)
    assert rel_l2(g.to_host(dr), o.vr_sor(other, u, ordr)) <= TOL
    g.set_vr_dirichlet(None)


@pytest.mark.parametrize("kind", ["simplex-2d", "mixed-2d", "simplex-3d"])
def test_vr_sor_simplex_meshes(require_gpu, kind):
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    print(f"\n[sor] {kind}", flush=True)
    if kind == "simplex-2d":
        m = mm.tri_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.1, seed=2)
    elif kind == "mixed-2d":
        m = mm.mixed_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.1, seed=2)
    else:
        m = mm.tet_grid(4, 3, 3, 1.0, 0.8, 0.6, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3, zmin=2, zmax=3),
                        jitter=0.1, seed=3)
    types = [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM]
    vals = np.tile([1.0, 0.3, 0.05, 0.0, 1 / 0.56 + 0.5 * 0.0925], (4, 1))
    o, g, ev = _pair(m, 2 if kind != "simplex-3d" else 1, types, vals, 1.0)
    Ao, bo = o.vr_coeffs()
    g.set_matrixAAInvB(Ao, bo)
    N, M = o.nCell, o.M
    u, rec = random_state(N, seed=5), random_rec(N, M, 5, seed=6)
    dr = g.to_device(rec)
    g.DoReconstructionSOR(dr, g.to_device(u))
    assert rel_l2(g.to_host(dr), o.vr_sor(rec, u, colour_scan_order(g.colors()))) <= TOL


@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_solver_loops_with_sor(require_gpu, kind):
    from dndsr_amd import mesh as mm
    from test_gpu_bc import _flow
    import pyoracle as po
    print(f"\n[sor steps] {kind}", flush=True)
    m0, types, vals = _bounded(kind)
    order = 1 if kind == "2d" else 2
    _, g0, _ = _pair(m0, order, types, vals, 1.0, sor=True)
    m = mm.colour_sorted(m0, g0.colors())
    o, g, ev = _pair(m, order, types, vals, 1.0, sor=True)
    ordr = colour_scan_order(g.colors())
    assert np.array_equal(ordr, np.arange(o.nCell))      # natural order is the colour order
    es = po.OracleVR.euler_settings(GAMMA, 1, 0, 1.0, 1.0)
    u0 = _flow(o, g, kind)
    du, dr = g.to_device(u0), g.BuildURec(5)
    ev.SSPRK3(du, dr, 3, 0.4, 2)
    uo, ro = o.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 3, 0.4, 2)
    assert rel_l2(g.to_host(du), uo) <= 1e-11
    assert rel_l2(g.to_host(dr), ro) <= 1e-10
    o_j = po.OracleVR(m, **dict(vr_settings(order), zoneDirWeight=g.settings["zoneDirWeight"]))
    o_j.set_bc(types, vals, GAMMA)
    assert rel_l2(o_j.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 3, 0.4, 2)[1], ro) > 1e-8   # SOR mattered
    dt = 5 * o.eval_dt(es, u0, 0.5, 1e100, False)[3]
    du, dr = g.to_device(u0), g.BuildURec(5)
    n, res, rel, hist = ev.ImplicitEulerStep(du, dr, dt, CFL=10.0, maxIter=3, minIter=3, tol=0.0)
    ui, ri, ni, *_ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), dt, CFL=10.0, maxIter=3,
                                           minIter=3, tol=0.0, scanOrder=ordr)
    assert n == ni == 3
    assert rel_l2(g.to_host(du), ui) <= 1e-11
    assert rel_l2(g.to_host(dr), ri) <= 1e-10
