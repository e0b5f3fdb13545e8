"""The benchmark configurations at their full sizes on the device.

* 2-D isentropic vortex, 64 x 64 periodic quads, p1, Roe, SSP-RK3 CFL 0.5, 100 steps:
  small enough for the CPU oracle to run the whole configuration, so the device's state
  after 100 steps is compared with the oracle's (relative L2 <= 1e-10).
* 3-D Taylor-Green vortex, 64^3 hexes, p2 (intOrder 5), HLLE+: the residual conserves
  mass, momentum and energy (sum V rhs = 0 to rounding) and 20 SSP-RK3 steps keep the
  totals (relative 1e-12), a property independent of the size.
* Laminar box, 1.97M jittered tetrahedra, p1 (intOrder 3), Roe, viscous: a uniform
  stream with the free stream imposed on every boundary zone is preserved (the
  reconstruction stays zero, the residual vanishes to rounding).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_isentropic_vortex_64_full_run_matches_oracle(require_gpu):
    import pyoracle as po
    from dndsr_amd import fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    print("\n[full size] IV 64^2 p1 Roe, 100 SSP-RK3 steps", flush=True)
    m = mm.periodic_quad_grid(64, 64, 10.0, 10.0)
    s = dict(maxOrder=1, intOrder=3, dirWeightScheme=0, geomWeightScheme=0, geomWeightBias=0.0,
             geomWeightPower=0.5, jacobiRelax=1.0)
    o = po.OracleVR(m, **s)
    g = VariationalReconstruction(m, s).construct()
    ub = fields.iv_background_for_mach(0.5)
    u0 = g.cell_average(lambda p: fields.isentropic_vortex(p, 5.0, 1.4, ub), 5)
    es = po.OracleVR.euler_settings(1.4, 1, 0, 1.0, 1.0)
    ev = EulerEvaluator(g, dict(riemannSolverType="Roe"))
    du, dr = g.to_device(u0), g.BuildURec(5)
    ev.SSPRK3(du, dr, 100, 0.5, 1)
    po.set_threads(8)
    uo, ro = o.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 100, 0.5, 1)
    assert rel_l2(g.to_host(du), uo) <= 1e-10
    assert rel_l2(g.to_host(dr), ro) <= 1e-8
    assert rel_l2(uo, u0) > 1e-3  # the vortex moved


def test_taylor_green_64_conservation(require_gpu):
    import torch
    from dndsr_amd import fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    print("\n[full size] TGV 64^3 p2 HLLE+", flush=True)
    L = 2 * np.pi
    m = mm.periodic_hex_grid(64, 64, 64, L, L, L)
    g = VariationalReconstruction(m, dict(maxOrder=2, intOrder=5)).construct()
    ev = EulerEvaluator(g, dict(gamma=1.4, riemannSolverType="HLLEP"))
    u = g.to_device(g.cell_average(lambda p: fields.taylor_green(p, 0.1, 1.4), 5))
    rec = g.BuildURec(5)
    vol = torch.from_numpy(g.cell_metrics()["vol"]).to(u.device)
    out = torch.empty_like(rec)
    g.DoReconstructionIter(rec, out, u)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, u, None, CFL=0.5)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, u, out, out)
    tot = (rhs * vol).sum(1)
    scale = (rhs.abs() * vol).sum(1)
    assert float((tot.abs() / scale).max()) <= 1e-13
    m0 = (u * vol).sum(1)
    ev.SSPRK3(u, rec, 20, 0.5, 1)
    m1 = (u * vol).sum(1)
    assert float(((m1 - m0).abs() / (u.abs() * vol).sum(1)).max()) <= 1e-12
    assert torch.isfinite(u).all()


def test_laminar_box_tetrahedra_free_stream(require_gpu):
    import torch
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    print("\n[full size] laminar box, 1.97M tetrahedra, free stream", flush=True)
    gamma = 1.4
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / gamma / (gamma - 1) + 0.5 * 0.25])
    types = [bc.BC_IN] * 4
    m = mm.bump_box_tet(69)
    g = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3, bcTypes=types)).construct()
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * 4), viscous=True,
                                idealGasProperty=dict(Rgas=1.0, muGas=0.5e-4, prGas=0.72, muModel=0)))
    assert g.nCell == 1971054
    u = g.to_device(np.tile(far, (g.nCell, 1)))
    rec = g.BuildURec(5)
    out = torch.empty_like(rec)
    g.DoReconstructionIter(rec, out, u)
    assert float(out.abs().max()) <= 1e-13
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, u, None, CFL=1.0)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, u, out, out)
    # flux scale: |F| ~ rho U^2 + p ~ 1, face area / volume ~ 1 / h ~ 70
    assert float(rhs.abs().max()) <= 1e-10


def test_taylor_green_64_matches_oracle(require_gpu):
    """The C2 configuration at its full size against the oracle: two Jacobi VR sweeps,
    EvaluateDt, EvaluateRHS on the swept coefficients and one SSP-RK3 step, all on the
    262144-cell p2 mesh (relative L2 <= 1e-12)."""
    import pyoracle as po
    import torch
    from dndsr_amd import fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    print("\n[full size] TGV 64^3 p2 HLLE+ against the oracle", flush=True)
    L = 2 * np.pi
    m = mm.periodic_hex_grid(64, 64, 64, L, L, L)
    s = dict(maxOrder=2, intOrder=5)
    g = VariationalReconstruction(m, s).construct()
    ev = EulerEvaluator(g, dict(gamma=1.4, riemannSolverType="HLLEP"))
    u0 = g.cell_average(lambda p: fields.taylor_green(p, 0.1, 1.4), 5)
    u = g.to_device(u0)
    rec = g.BuildURec(5)
    out = torch.empty_like(rec)
    g.DoReconstructionIter(rec, out, u)
    g.DoReconstructionIter(out, rec, u)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, u, None, CFL=0.5, UseLocaldt=False)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, u, rec, rec)
    rec_d, rhs_d, dlc_d = g.to_host(rec), g.to_host(rhs), ev.deltaLambdaCell.cpu().numpy()
    u1 = u.clone()
    r1 = g.BuildURec(5)
    ev.SSPRK3(u1, r1, 1, 0.5, 1)
    u1_d = g.to_host(u1)

    po.set_threads(min(16, os.cpu_count() or 1))
    o = po.OracleVR(m, **s)
    es = o.euler_settings(1.4, 3, 0, 1.0, 1.0)
    ro = o.vr_iter(o.vr_iter(np.zeros((o.nCell, 5, o.M)), u0), u0)
    assert rel_l2(rec_d, ro) <= 1e-12
    _, _, dlc_o, _ = o.eval_dt(es, u0, 0.5, 1e100, False)
    assert rel_l2(dlc_d, dlc_o) <= 1e-13
    rhs_o = o.eval_rhs(es, u0, ro, dlc_o)
    assert rel_l2(rhs_d, rhs_o) <= 1e-12
    u1_o, _ = o.ssprk3(es, u0, np.zeros((o.nCell, 5, o.M)), 1, 0.5, 1)
    assert rel_l2(u1_d, u1_o) <= 1e-12
    assert rel_l2(u1_o, u0) > 1e-6  # the step changed the state


def _laminar_box(n, Re=1e4):
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    gamma = 1.4
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / gamma / (gamma - 1) + 0.5 * 0.25])  # Mach 0.5, a = 1
    types = [bc.BC_IN, bc.BC_IN, bc.BC_WALL, bc.BC_IN]
    gas = dict(Rgas=1.0, muGas=0.5 / Re, prGas=0.72, muModel=0)
    return mm.bump_box_tet(n), gamma, far, types, gas


def _perturbed_stream(cent, far, gamma):
    """The free stream with a smooth 5 % perturbation of density, velocity and pressure."""
    x, y, z = cent[:, 0], cent[:, 1], cent[:, 2]
    rho = far[0] * (1 + 0.05 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z))
    vel = np.stack([0.5 * (1 + 0.05 * np.cos(np.pi * y)), 0.02 * np.sin(np.pi * x) * np.sin(np.pi * y),
                    0.02 * np.cos(2 * np.pi * z) * np.sin(np.pi * y)], 1)
    p = 1 / gamma * (1 + 0.05 * np.cos(2 * np.pi * x) * np.sin(np.pi * y))
    E = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)
    return np.concatenate([rho[:, None], rho[:, None] * vel, E[:, None]], 1)


def test_laminar_box_tetrahedra_matches_oracle(require_gpu):
    """The C3 configuration at its full size (1.97M tetrahedra, viscous, no-slip wall,
    free-stream boundaries) on a perturbed stream against the oracle: a VR sweep,
    EvaluateDt, the viscous residual, and two implicit pseudo-time iterations with FP64
    sweeps, the oracle scanning cells in the device's colour order."""
    import pyoracle as po
    import torch
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    from helpers import RS, colour_scan_order
    print("\n[full size] laminar box, 1.97M tetrahedra, perturbed stream against the oracle", flush=True)
    m, gamma, far, types, gas = _laminar_box(69)
    vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    g = VariationalReconstruction(m, vs).construct()
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * 4), viscous=True, idealGasProperty=gas))
    u0 = _perturbed_stream(g.cell_metrics()["cent"], far, gamma)
    u = g.to_device(u0)
    rec = g.BuildURec(5)
    out = torch.empty_like(rec)
    g.DoReconstructionIter(rec, out, u)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, u, None, CFL=1000.0)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, u, out, out)
    rec_d, rhs_d, dt_d = g.to_host(out), g.to_host(rhs), dt.cpu().numpy()
    order = colour_scan_order(g.colors())
    u2, r2 = u.clone(), g.BuildURec(5)
    n2, _, _, _ = ev.ImplicitEulerStep(u2, r2, 1e100, CFL=1000.0, sgsIter=3, maxIter=2, minIter=2, tol=0.0)
    u2_d = g.to_host(u2)
    del g, ev, u, rec, out, rhs, u2, r2

    po.set_threads(min(16, os.cpu_count() or 1))
    o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o.set_bc(types, np.stack([far] * 4), gamma)
    es = po.OracleVR.euler_settings(gamma, RS["Roe"], 0, 1.0, 1.0, viscous=1, **gas)
    ro = o.vr_iter(np.zeros((o.nCell, 5, o.M)), u0)
    assert rel_l2(rec_d, ro) <= 1e-12
    dt_o, _, _, _ = o.eval_dt(es, u0, 1000.0, 1e100, True)
    assert rel_l2(dt_d, dt_o) <= 1e-13
    rhs_o = o.eval_rhs(es, u0, ro)
    assert rel_l2(rhs_d, rhs_o) <= 1e-12
    u2_o, _, n2_o, _, _, _ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), 1e100, CFL=1000.0,
                                                    sgsIter=3, maxIter=2, minIter=2, tol=0.0, scanOrder=order)
    assert n2 == n2_o == 2
    assert rel_l2(u2_d, u2_o) <= 1e-11
    assert rel_l2(u2_o, u0) > 1e-4  # the iterations moved the state


def test_laminar_box_reduced_converged_matches_oracle(require_gpu):
    """The C3 configuration (same Re, Mach, boundaries, p1, Roe, CFL 1000, three SGS
    sweeps) on a 10^3 x 6 = 6000-tetrahedron box, converged to a density residual of
    1e-12 of its maximum with FP64 sweeps, with FP32 sweeps and in the oracle: both
    device states match the oracle's converged state to <= 1e-10."""
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    from helpers import RS
    print("\n[reduced] laminar box, 6000 tetrahedra, converged fp64 / fp32 sweeps / oracle", flush=True)
    m, gamma, far, types, gas = _laminar_box(10)
    vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    g = VariationalReconstruction(m, vs).construct()
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * 4), viscous=True, idealGasProperty=gas))
    u0 = np.tile(far, (g.nCell, 1))
    o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o.set_bc(types, np.stack([far] * 4), gamma)
    es = po.OracleVR.euler_settings(gamma, RS["Roe"], 0, 1.0, 1.0, viscous=1, **gas)
    uo, _, no, _, relo, _ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), 1e100, CFL=1000.0,
                                                   sgsIter=3, maxIter=5000, minIter=5, tol=1e-12)
    assert relo[0] < 1e-12, (no, relo)
    assert rel_l2(uo, u0) > 1e-3  # the wall and the bump changed the flow
    for fp32 in (False, True):
        u, rec = g.to_device(u0), g.BuildURec(5)
        n, _, rel, _ = ev.ImplicitEulerStep(u, rec, 1e100, CFL=1000.0, sgsIter=3, maxIter=5000, minIter=5,
                                            tol=1e-12, fp32Sweeps=fp32)
        assert rel[0] < 1e-12, (fp32, n, rel)
        assert rel_l2(g.to_host(u), uo) <= 1e-10, fp32
