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
