"""The steady discrete adjoint on the device (no counterpart in the reference).

* The transposed scalar LU-SGS sweep (dndsr_update_sgs_transpose), the adjoint's
  preconditioner, against the oracle's sweep in the device's colour order (1e-12); the
  oracle's sweep is pinned as a sweep of the assembled transposed LU-SGS matrix on the CPU.
  The symmetric multi-sweep call (dndsr_sgs_transpose_symmetric, scalar and block) equals
  the chain of single sweeps bitwise and the oracle's chain to 1e-12.
* The gradient of the wall drag of a converged viscous bump-channel flow with respect to the
  five far-field values: SteadyAdjoint's multipliers against central differences of the drag
  of oracle steady states converged to 1e-13 at perturbed far-field values (1e-6 relative).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RS, colour_scan_order, random_state, rel_l2

pytestmark = pytest.mark.gpu
GAMMA = 1.4


def _sweep_case(kind):
    import pyoracle as po
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    if kind == "periodic":
        m = mm.periodic_hex_grid(5, 4, 4, 1.0, 0.8, 0.8, jitter=0.1, seed=3)
        o = po.OracleVR(m, maxOrder=1, intOrder=3)
        g = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3)).construct()
        ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe"))
    else:
        types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
        m = mm.box_quad_grid(7, 6, 7.0, 6.0, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.15, seed=2)
        vals = np.tile(random_state(1, seed=9)[0], (3, 1))
        o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
        o.set_bc(types, vals, GAMMA)
        g = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3, bcTypes=types)).construct()
        ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=vals))
    return o, g, ev


def _check_symmetric(ev, g, du, jd, rhs, gsUpdate, oracle_sweep):
    """SGSTransposeSymmetric (one call, colour passes that repeat the previous sweep's last
    one skipped under gsUpdate) equals the chain of single transposed sweeps from zero to the
    last bit, and the oracle's chain in the device's colour order to 1e-12."""
    import torch
    N = rhs.shape[0]
    r = g.to_device(rhs)
    for nS in (1, 3):
        a, b = torch.zeros_like(r), torch.zeros_like(r)
        y = np.zeros((N, 5))
        for _ in range(nS):
            for fw in (True, False):
                ev.UpdateSGSTranspose(1.0, r, du, a, b, jd, fw, gsUpdate)
                a, b = b, a
                y = oracle_sweep(y, fw)
        out = ev.SGSTransposeSymmetric(1.0, r, du, jd, nS, gsUpdate)
        assert torch.equal(out, a), (nS, float((out - a).abs().max()))
        assert rel_l2(g.to_host(out), y) <= 1e-12, nS


@pytest.mark.parametrize("kind", ["periodic", "bounded"])
@pytest.mark.parametrize("gsUpdate", [0, 1])
def test_transposed_sweep_matches_oracle(require_gpu, kind, gsUpdate):
    import pyoracle as po
    import torch
    o, g, ev = _sweep_case(kind)
    N = o.nCell
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"])
    u = random_state(N, seed=4)
    du = g.to_device(u)
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=2.0)
    jd = torch.empty_like(dTau)
    ev.LUSGSMatrixInit(jd, None, dTau, 0.7, 1.0)
    dTau_o, lam_o, _, _ = o.eval_dt(es, u, 2.0, 1e100, True)
    jd_o = o.lusgs_init(dTau_o, 0.7, 1.0, lam_o)
    assert rel_l2(jd.cpu().numpy(), jd_o) <= 1e-13
    rng = np.random.default_rng(8)
    rhs, y0 = rng.normal(size=(N, 5)), rng.normal(size=(N, 5))
    order = colour_scan_order(g.colors())
    for forward in (True, False):
        for zero in (False, True):
            out = torch.zeros_like(du)
            s = ev.UpdateSGSTranspose(1.0, g.to_device(rhs), du, g.to_device(y0), out, jd, forward, gsUpdate,
                                      uIncIsZero=zero)
            ref, s_o = o.update_sgs_transpose(es, 1.0, rhs, u, y0, np.zeros((N, 5)), jd_o, lam_o, forward, gsUpdate,
                                              zero, scanOrder=order)
            assert rel_l2(g.to_host(out), ref) <= 1e-12, (forward, zero)
            np.testing.assert_allclose(s, s_o, rtol=1e-11)
    _check_symmetric(ev, g, du, jd, rhs, gsUpdate, lambda y, fw: o.update_sgs_transpose(
        es, 1.0, rhs, u, y, np.zeros((N, 5)), jd_o, lam_o, fw, gsUpdate, False, scanOrder=order)[0])


@pytest.mark.parametrize("kind", ["periodic", "bounded"])
@pytest.mark.parametrize("gsUpdate", [0, 1])
def test_transposed_block_sweep_matches_oracle(require_gpu, kind, gsUpdate):
    """The transposed Roe-Jacobian block sweep (dndsr_update_sgs_block_transpose) against the
    oracle's in the device's colour order, viscous, with the face eigenvalues of EvaluateRHS."""
    import pyoracle as po
    import torch
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    gas = dict(Rgas=1.0, muGas=0.05, prGas=0.72, muModel=0)
    if kind == "periodic":
        m = mm.periodic_hex_grid(5, 4, 4, 1.0, 0.8, 0.8, jitter=0.1, seed=3)
        types = vals = None
        o = po.OracleVR(m, maxOrder=1, intOrder=3)
        vs = dict(maxOrder=1, intOrder=3)
    else:
        types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
        m = mm.box_quad_grid(7, 6, 7.0, 6.0, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.15, seed=2)
        vals = np.tile(random_state(1, seed=9)[0], (3, 1))
        o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
        o.set_bc(types, vals, GAMMA)
        vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    g = VariationalReconstruction(m, vs).construct()
    st = dict(gamma=GAMMA, riemannSolverType="Roe", viscous=True, idealGasProperty=gas, useRoeJacobian=True)
    if types is not None:
        st.update(bcTypes=types, bcValues=vals)
    ev = EulerEvaluator(g, st)
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], viscous=1, useRoeJacobian=1, **gas)
    N = o.nCell
    u = random_state(N, seed=4)
    rec = 0.05 * np.random.default_rng(5).normal(size=(N, 5, o.M))
    du, dr = g.to_device(u), g.to_device(rec)
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=2.0)
    ev.EvaluateRHS(g.BuildUDof(5), None, du, dr)
    jd = ev.BuildJacobianDiagBlock()
    ev.LUSGSMatrixInit(jd, None, dTau, 0.7, 1.0, du, dr)
    dTau_o, lam_o, dlc, _ = o.eval_dt(es, u, 2.0, 1e100, True)
    _, lr_o = o.eval_rhs(es, u, rec, dlc, roeLambdas=True)
    _, ji_o = o.lusgs_init_block(es, dTau_o, 0.7, 1.0, u, lam_o, lr_o)
    rng = np.random.default_rng(8)
    rhs, y0 = rng.normal(size=(N, 5)), rng.normal(size=(N, 5))
    order = colour_scan_order(g.colors())
    for forward in (True, False):
        for zero in (False, True):
            out = torch.zeros_like(du)
            s = ev.UpdateSGSTranspose(1.0, g.to_device(rhs), du, g.to_device(y0), out, jd, forward, gsUpdate,
                                      uIncIsZero=zero)
            ref, s_o = o.update_sgs_block_transpose(es, 1.0, rhs, u, y0, np.zeros((N, 5)), ji_o, lam_o, lr_o,
                                                    forward, gsUpdate, zero, scanOrder=order)
            assert rel_l2(g.to_host(out), ref) <= 1e-12, (forward, zero)
            np.testing.assert_allclose(s, s_o, rtol=1e-11)
    _check_symmetric(ev, g, du, jd, rhs, gsUpdate, lambda y, fw: o.update_sgs_block_transpose(
        es, 1.0, rhs, u, y, np.zeros((N, 5)), ji_o, lam_o, lr_o, fw, gsUpdate, False, scanOrder=order)[0])


def _bump_channel():
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    types = [bc.BC_WALL, bc.BC_FAR, bc.BC_SYM]
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / GAMMA / (GAMMA - 1) + 0.125])

    def bump(P):
        Q = P.copy()
        Q[:, 1] += 0.05 * np.sin(np.pi * np.clip(P[:, 0], 0, 1)) ** 2 * (1 - P[:, 1])
        return Q

    m = mm.box_quad_grid(16, 8, 2.0, 1.0, x0=-0.5, zones=dict(xmin=1, xmax=1, ymin=0, ymax=2), jitter=0.05,
                         seed=1, node_map=bump)
    gas = dict(Rgas=1.0, muGas=1e-2, prGas=0.72, muModel=0)
    return m, types, far, gas


STEADY = dict(CFL=20.0, sgsIter=2, maxIter=6000, minIter=5, tol=1e-13)


def _oracle_drag(o, es, types, far, u0, farValue):
    vals = np.stack([far] * 3)
    vals[1] = farValue
    o.set_bc(types, vals, GAMMA)
    uo, ro, _, _, relo, _ = o.implicit_euler_step(es, u0, np.zeros((o.nCell, 5, o.M)), 1e100, **STEADY)
    assert relo[0] < 1e-12
    _, _, dlc, _ = o.eval_dt(es, uo, 1.0, 1e100, True)
    return o.zone_force(es, uo, ro, dlc, 0)[0], uo


@pytest.mark.parametrize("subs2nd", [0, 1])
def test_steady_adjoint_drag_gradient_in_far_field_values(require_gpu, subs2nd):
    import pyoracle as po
    import torch
    from dndsr_amd import bc
    from dndsr_amd.adjoint import SteadyAdjoint, zone_force, zone_force_vjp
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    m, types, far, gas = _bump_channel()
    vs = dict(maxOrder=1, intOrder=3, subs2ndOrder=subs2nd)
    o = po.OracleVR(m, **vs, zoneDirWeight=bc.euler_dir_weights(types, 1))
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], viscous=1, **gas)
    u0 = np.tile(far, (o.nCell, 1))
    J0, uo = _oracle_drag(o, es, types, far, u0, far)
    h = 1e-5
    g_fd = np.zeros(5)
    for k in range(5):
        p, q = far.copy(), far.copy()
        p[k] += h
        q[k] -= h
        g_fd[k] = (_oracle_drag(o, es, types, far, u0, p)[0] - _oracle_drag(o, es, types, far, u0, q)[0]) / (2 * h)

    g = VariationalReconstruction(m, dict(vs, bcTypes=types)).construct()
    ev = EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=np.stack([far] * 3),
                                viscous=True, idealGasProperty=gas))
    du, dr = g.to_device(u0), g.BuildURec(5)
    n, _, rel, _ = ev.ImplicitEulerStep(du, dr, 1e100, **STEADY)
    assert rel[0] < 1e-12, n
    assert rel_l2(g.to_host(du), uo) <= 1e-10
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=1.0)
    assert abs(zone_force(ev, du, dr, 0)[0] - J0) <= 1e-10 * abs(J0)

    adj = SteadyAdjoint(ev, CFL=20.0, sgsIter=1)
    uB, rB = zone_force_vjp(ev, du, dr, 0, [1.0, 0.0, 0.0])

    def drag(ev_, u, rec, dlc):
        return zone_force(ev_, u, rec, 0, dlc)[0]

    lams = []
    for method in ("fixed-point", "gmres"):
        lam, mu, nIt = adj.solve(du, dr, uB, rB, tol=1e-11, maxIter=4000, method=method)
        assert nIt < 4000, method
        lams.append(lam)
        grad = adj.bc_value_gradient(du, dr, lam, mu, drag, zones=[1])[0]
        assert np.abs(grad - g_fd).max() <= 1e-6 * np.abs(g_fd).max(), (method, grad, g_fd)
    # both solvers land on the same multipliers
    assert float(torch.linalg.vector_norm(lams[0] - lams[1])) <= 1e-8 * float(torch.linalg.vector_norm(lams[1]))
    # the boundary values are restored
    assert abs(zone_force(ev, du, dr, 0)[0] - J0) <= 1e-10 * abs(J0)
    assert torch.isfinite(lam).all()


def test_steady_adjoint_refuses_fixed_delta_lambda_forms(require_gpu):
    from dndsr_amd.adjoint import SteadyAdjoint
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    from dndsr_amd import mesh as mm
    m = mm.periodic_quad_grid(4, 4, 4.0, 4.0)
    g = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3)).construct()
    with pytest.raises(ValueError, match="deltaLambdaCell"):
        SteadyAdjoint(EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="HLLEP")))
    SteadyAdjoint(EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="HLLEP", rsFixScale=0.0)))


def _c4_reduced():
    """The 2-D adjoint workload (Hicks-Henne bump, Mach 0.7, mixed cells) at 192 cells and
    Re 200, where the oracle converges a steady state in about 600 block LU-SGS iterations."""
    from dndsr_amd import bc
    from dndsr_amd.shape import HicksHenneBumpChannel
    ch = HicksHenneBumpChannel(16, 8)
    alpha = np.random.default_rng(20261203).uniform(-0.01, 0.01, 8)
    vals, gas = ch.flow(M=0.7, Re=200.0)
    types = [bc.bc_code(t) for t in ch.BC_TYPES]
    return ch, alpha, vals, gas, types


C4_STEADY = dict(CFL=500.0, sgsIter=4, maxIter=20000, minIter=5, tol=1e-13)


def _oracle_c4_drag(ch, a, vals, gas, types):
    import pyoracle as po
    from dndsr_amd import bc
    o = po.OracleVR(ch.mesh(a), maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o.set_bc(types, vals, GAMMA)
    es = po.OracleVR.euler_settings(GAMMA, RS["Roe"], viscous=1, useRoeJacobian=1,
                                    **{k: gas[k] for k in ("Rgas", "muGas", "prGas", "muModel")})
    uo, ro, _, _, relo, _ = o.implicit_euler_step(es, np.tile(vals[1], (o.nCell, 1)), np.zeros((o.nCell, 5, o.M)),
                                                  1e100, **C4_STEADY)
    assert relo[0] < 1e-12
    _, _, dlc, _ = o.eval_dt(es, uo, 1.0, 1e100, True)
    return o.zone_force(es, uo, ro, dlc, 0)[0], uo


@pytest.mark.parametrize("method,block", [("gmres", True), ("fixed-point", True), ("gmres", False)])
def test_shape_gradient_of_drag_matches_oracle_fd(require_gpu, method, block):
    """Hicks-Henne shape gradient of the wall drag: adjoint multipliers on the device, the
    Lagrangian differenced in each parameter on moved meshes (step 1e-6), against central
    differences (step 1e-6) of the drag of oracle steady states on the moved meshes."""
    from dndsr_amd.adjoint import SteadyAdjoint, zone_force, zone_force_vjp
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    from dndsr_amd.shape import shape_gradient
    ch, alpha, vals, gas, types = _c4_reduced()
    J0, uo = _oracle_c4_drag(ch, alpha, vals, gas, types)
    h = 1e-6
    g_fd = np.zeros(8)
    for k in range(8):
        e = np.zeros(8)
        e[k] = h
        g_fd[k] = (_oracle_c4_drag(ch, alpha + e, vals, gas, types)[0]
                   - _oracle_c4_drag(ch, alpha - e, vals, gas, types)[0]) / (2 * h)

    def build(a, roeJ=False):
        g = VariationalReconstruction(ch.mesh(a), dict(maxOrder=1, intOrder=3, bcTypes=types)).construct()
        return EulerEvaluator(g, dict(gamma=GAMMA, riemannSolverType="Roe", bcTypes=types, bcValues=vals,
                                      viscous=True, idealGasProperty=gas, useRoeJacobian=roeJ))

    ev = build(alpha, roeJ=True)
    g = ev.vfv
    du, dr = g.to_device(np.tile(vals[1], (g.nCell, 1))), g.BuildURec(5)
    n, _, rel, _ = ev.ImplicitEulerStep(du, dr, 1e100, **C4_STEADY)
    assert rel[0] < 1e-12, n
    assert rel_l2(g.to_host(du), uo) <= 1e-10
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=1.0)
    assert abs(zone_force(ev, du, dr, 0)[0] - J0) <= 1e-10 * abs(J0)
    adj = SteadyAdjoint(ev, CFL=50.0, sgsIter=2, block=block)
    uB, rB = zone_force_vjp(ev, du, dr, 0, [1.0, 0.0, 0.0])
    lam, mu, nIt = adj.solve(du, dr, uB, rB, tol=1e-12, maxIter=6000, method=method)
    assert nIt < 6000, adj.history[-3:]

    def drag(ev_, u, rec, dlc):
        return zone_force(ev_, u, rec, 0, dlc)[0]

    gS = shape_gradient(adj, du, dr, lam, mu, drag, build, alpha, h=1e-6)
    err = np.abs(gS - g_fd).max() / np.abs(g_fd).max()
    print(f"shape gradient ({method}, block {block}, {nIt} iterations): max rel error vs oracle FD {err:.3e}")
    assert err <= 1e-6, (gS, g_fd)


def test_symmetric_transposed_sweeps_reject_bad_arguments(require_gpu):
    """dndsr_sgs_transpose_symmetric: exactly one of the scalar and block Jacobians, nSweep >= 1,
    distinct work / out / rhs arrays, lamRoe with the block Jacobian; each refusal names its
    cause and leaves out untouched."""
    import ctypes as C

    import torch
    from dndsr_amd import _lib
    o, g, ev = _sweep_case("bounded")
    N = o.nCell
    du = g.to_device(random_state(N, seed=4))
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, du, None, CFL=2.0)
    jd = torch.empty_like(dTau)
    ev.LUSGSMatrixInit(jd, None, dTau, 0.7, 1.0)
    rhs = g.to_device(np.random.default_rng(3).normal(size=(N, 5)))
    out, work = torch.full_like(rhs, 7.0), torch.zeros_like(rhs)
    L = _lib.lib()
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731

    def call(jdiag, jinv, lamRoe, nSweep, w, o_):
        return L.dndsr_sgs_transpose_symmetric(ev.ctx.h, C.byref(ev.es), 1.0, p(rhs), p(du), p(jdiag), p(jinv),
                                               p(ev.lambdaFace), p(lamRoe), nSweep, 1, p(w), p(o_))

    cases = [((jd, jd, ev.lambdaFaceRoe, 2, work, out), "exactly one of jdiag"),
             ((None, None, None, 2, work, out), "exactly one of jdiag"),
             ((jd, None, None, 0, work, out), "nSweep must be at least 1"),
             ((jd, None, None, 2, out, out), "different arrays"),
             ((None, jd, None, 2, work, out), "lamRoe")]
    for args, msg in cases:
        assert call(*args) != 0
        assert msg in L.dndsr_last_error().decode(), (msg, L.dndsr_last_error())
    assert bool((out == 7.0).all())
    assert call(jd, None, None, 2, work, out) == 0
