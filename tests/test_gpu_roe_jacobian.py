"""The Roe-Jacobian block LU-SGS (useRoeJacobian) on the device against the oracle:
the face eigenvalues EvaluateRHS keeps (EulerEvaluator_EvaluateRHS.hxx:527-544), the 5x5
JDiag blocks and their inverses (LUSGSMatrixInit, EulerEvaluator.hxx:37-159; GetInvert,
EulerJacobian.hpp:153-190), the block sweeps in the device's colour order (UpdateSGS,
EulerEvaluator.hxx:550-693; FP64 to 1e-12, FP32 to single precision) and the implicit step with
the block Jacobian (fixed iterations to 1e-11; converged FP64 and FP32-sweep states within 1e-10
of the oracle's)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, colour_scan_order, random_rec, random_state, rel_l2

pytestmark = pytest.mark.gpu
GAMMA = 1.4


def _case(kind):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    types = vals = None
    if kind == "2d-periodic":
        m = mm.periodic_quad_grid(7, 6, 7.0, 6.0, jitter=0.15, seed=2)
        s = dict(maxOrder=1, intOrder=3)
    elif kind == "3d-periodic":
        m = mm.periodic_hex_grid(5, 4, 4, 1.0, 0.8, 0.8, jitter=0.1, seed=3)
        s = dict(maxOrder=2, intOrder=5)
    else:  # bounded quads: no-slip wall, symmetry, far field, inflow, outflow
        types = [bc.BC_WALL, bc.BC_SYM, bc.BC_FAR, bc.BC_IN, bc.BC_OUT]
        m = mm.box_quad_grid(7, 6, 7.0, 6.0, zones=dict(xmin=3, xmax=4, ymin=0, ymax=2), jitter=0.15, seed=2)
        up = np.where(m.faceZone == 2)[0]
        m.faceZone[up[: len(up) // 2]] = 1
        m.nBCZone = 5
        vals = np.tile(random_state(1, seed=9)[0], (5, 1))
        s = dict(maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o = po.OracleVR(m, **s)
    if types is not None:
        o.set_bc(types, vals, GAMMA)
    g = VariationalReconstruction(m, s).construct()
    return m, o, g, types, vals


def _ev(g, types, vals, rs="Roe", batch=0):
    from dndsr_amd.euler import EulerEvaluator
    st = dict(gamma=GAMMA, riemannSolverType=rs, rsMeanValueEig=batch, useRoeJacobian=True)
    if types is not None:
        st.update(bcTypes=types, bcValues=vals)
    return EulerEvaluator(g, st)


KINDS = ["2d-periodic", "3d-periodic", "bounded"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("rs,batch", [("Roe", 0), ("Roe_M1", 1), ("Roe_M6", 0), ("HLLEP", 0)])
def test_face_eigenvalues(require_gpu, kind, rs, batch):
    import pyoracle as po
    m, o, g, types, vals = _case(kind)
    ev = _ev(g, types, vals, rs, batch)
    es = po.OracleVR.euler_settings(GAMMA, RS[rs], batch, useRoeJacobian=1)
    u, rec = random_state(o.nCell, seed=4), random_rec(o.nCell, o.M, seed=5)
    du, dr = g.to_device(u), g.to_device(rec)
    dt = g.BuildUDof(1)[0]
    ev.EvaluateDt(dt, du, None, CFL=1.0)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, du, dr, dr)
    _, _, dlc, _ = o.eval_dt(es, u, 1.0, 1e100, True)
    rhs_o, lr_o = o.eval_rhs(es, u, rec, dlc, roeLambdas=True)
    assert rel_l2(g.to_host(rhs), rhs_o) <= 1e-12
    assert rel_l2(ev.lambdaFaceRoe.cpu().numpy().T, lr_o) <= 1e-12


def _blocks(kind):
    import pyoracle as po
    m, o, g, types, vals = _case(kind)
    ev = _ev(g, types, vals)
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], useRoeJacobian=1)
    u, rec = random_state(o.nCell, seed=4), random_rec(o.nCell, o.M, seed=5)
    du, dr = g.to_device(u), g.to_device(rec)
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=2.0)
    rhs = g.BuildUDof(5)
    ev.EvaluateRHS(rhs, None, du, dr, dr)
    jd = ev.BuildJacobianDiagBlock()
    ev.LUSGSMatrixInit(jd, None, dTau, 0.7, 1.0, du, dr)
    dTau_o, lam_o, dlc, _ = o.eval_dt(es, u, 2.0, 1e100, True)
    _, lr_o = o.eval_rhs(es, u, rec, dlc, roeLambdas=True)
    jd_o, ji_o = o.lusgs_init_block(es, dTau_o, 0.7, 1.0, u, lam_o, lr_o)
    return m, o, g, ev, es, u, du, jd, jd_o, ji_o, lam_o, lr_o


@pytest.mark.parametrize("kind", KINDS)
def test_block_init(require_gpu, kind):
    m, o, g, ev, es, u, du, jd, jd_o, ji_o, _, _ = _blocks(kind)
    N = o.nCell
    assert rel_l2(g.to_host(jd.value), jd_o.reshape(N, 25)) <= 1e-12
    assert rel_l2(g.to_host(jd.invert), ji_o.reshape(N, 25)) <= 1e-12


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("gsUpdate", [0, 1])
def test_block_sweep(require_gpu, kind, gsUpdate):
    import torch
    m, o, g, ev, es, u, du, jd, jd_o, ji_o, lam_o, lr_o = _blocks(kind)
    N = o.nCell
    rng = np.random.default_rng(8)
    rhs, x0 = rng.normal(size=(N, 5)), 0.1 * rng.normal(size=(N, 5))
    order = colour_scan_order(g.colors())
    drhs, dx0 = g.to_device(rhs), g.to_device(x0)
    for forward in (True, False):
        out = torch.zeros_like(dx0)
        s = ev.UpdateSGS(1.0, 0.0, drhs, du, dx0, out, jd, forward, gsUpdate)
        ref, s_o = o.update_sgs_block(es, 1.0, rhs, u, x0, np.zeros((N, 5)), ji_o, lam_o, lr_o, forward, gsUpdate,
                                      False, scanOrder=order)
        assert rel_l2(g.to_host(out), ref) <= 1e-12
        np.testing.assert_allclose(s, s_o, rtol=1e-11)
        # FP32 sweep
        jf = jd.to(torch.float32)
        outf = torch.zeros_like(dx0, dtype=torch.float32)
        ev.UpdateSGS(1.0, 0.0, drhs.float(), du.float(), dx0.float(), outf, jf, forward, gsUpdate,
                     lambdaFace=ev.lambdaFace.float())
        assert rel_l2(g.to_host(outf.double()), ref) <= 2e-5


@pytest.mark.parametrize("kind", ["2d-periodic", "bounded"])
def test_implicit_fixed_iterations(require_gpu, kind):
    import pyoracle as po
    m, o, g, types, vals = _case(kind)
    ev = _ev(g, types, vals)
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], useRoeJacobian=1)
    u0 = random_state(o.nCell, seed=6)
    order = colour_scan_order(g.colors())
    dt = 0.5
    for nIt in (1, 3):
        du, dr = g.to_device(u0), g.BuildURec(5)
        n, res, rel, hist = ev.ImplicitEulerStep(du, dr, dt, CFL=5.0, sgsIter=1, maxIter=nIt, minIter=nIt, tol=0.0)
        uo, ro, no, reso, relo, histo = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), dt, CFL=5.0,
                                                              sgsIter=1, maxIter=nIt, minIter=nIt, tol=0.0,
                                                              scanOrder=order)
        assert n == no == nIt
        assert rel_l2(g.to_host(du), uo) <= 1e-11
        np.testing.assert_allclose(hist, histo, rtol=1e-9)


def test_implicit_converged_fp64_fp32_oracle(require_gpu):
    """The bounded case marched to a steady state (infinite physical step) with the Roe-Jacobian
    block LU-SGS: FP64 and FP32 sweeps both land within 1e-10 of the oracle's converged state."""
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / GAMMA / (GAMMA - 1) + 0.125])

    def bump(P):
        Q = P.copy()
        Q[:, 1] += 0.05 * np.sin(np.pi * np.clip(P[:, 0], 0, 1)) ** 2 * (1 - P[:, 1])
        return Q

    m = mm.box_quad_grid(16, 8, 2.0, 1.0, x0=-0.5, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.05,
                         seed=1, node_map=bump)
    zw = bc.euler_dir_weights(types, 1)
    o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=zw)
    o.set_bc(types, np.stack([far] * 3), GAMMA)
    gas = dict(Rgas=1.0, muGas=1e-2, prGas=0.72, muModel=0)
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], viscous=1, useRoeJacobian=1, **gas)
    u0 = np.tile(far, (o.nCell, 1))
    uo, _, no, _, relo, _ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), 1e100, CFL=50.0, sgsIter=2,
                                                   maxIter=3000, minIter=5, tol=1e-11)
    assert relo[0] < 1e-11
    g = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3, bcTypes=types)).construct()
    ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=np.stack([far] * 3),
                                viscous=True, idealGasProperty=gas, useRoeJacobian=True))
    for fp32 in (False, True):
        du, dr = g.to_device(u0), g.BuildURec(5)
        n, _, rel, _ = ev.ImplicitEulerStep(du, dr, 1e100, CFL=50.0, sgsIter=2, maxIter=3000, minIter=5, tol=1e-11,
                                            fp32Sweeps=fp32)
        assert rel[0] < 1e-11, (fp32, n)
        assert rel_l2(g.to_host(du), uo) <= 1e-10, fp32
