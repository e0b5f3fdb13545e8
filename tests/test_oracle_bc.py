"""CPU pins of the oracle's boundary conditions: ghost states (generateBoundaryValue,
src/Euler/EulerEvaluator_EvaluateDt.hxx:1743-2390), the boundary term of the VR sweep
(GetBoundaryRHS, src/CFV/VariationalReconstruction_Reconstruction.hxx:19-96) and
boundary faces in EvaluateRHS (EulerEvaluator_EvaluateRHS.hxx:194-207,336-345).

Pinned by what the mathematics fixes, not by retyping formulas:
* ghost states: the characteristic cases of the far-field condition reduce to closed
  forms (supersonic in/out: the exterior / interior state; subsonic in: exterior density
  and velocity with interior pressure; subsonic out: interior density and velocity with
  exterior pressure); a slip wall reverses the normal momentum only;
* VR with Dirichlet boundary values: a constant is reproduced for any boundary weights
  (the reference's Wall/const test, test/cpp/CFV/test_Reconstruction.cpp:455-480, 1e-12),
  and with value-only boundary weights (the weights EulerEvaluator sets) every
  polynomial of degree <= p is reproduced -- which fails if the boundary term and the
  boundary part of A disagree by any factor (checked: all-order weights do not
  reproduce a linear field);
* EvaluateRHS: a uniform flow is preserved by far-field and tangential slip walls, and
  the residual of a mirror-symmetric state next to a symmetry plane equals the residual
  of the same cells in the mirrored full domain (ghost = mirror image of the point state,
  mean ghost for the Roe average).
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import bc
from dndsr_amd import mesh as mm
from helpers import RS, rel_l2

GAMMA = 1.4


def _cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])


def _prim(U):
    rho = U[0]
    v = U[1:4] / rho
    return rho, v, (GAMMA - 1) * (U[4] - 0.5 * rho * v @ v)


# ---------------------------------------------------------------------------- ghost states
N_ = np.array([0.6, 0.8, 0.0])
FAR = _cons(1.3, [0.2, -0.1, 0.05], 0.9)


def test_far_field_characteristic_cases():
    a = np.sqrt(GAMMA * 1.0 / 1.0)
    for un, case in [(1.5 * a, "sup_out"), (0.4 * a, "sub_out"), (-0.4 * a, "sub_in"), (-1.5 * a, "sup_in")]:
        tang = np.array([-0.8, 0.6, 0.0]) * 0.3
        UL = _cons(1.0, un * N_ + tang, 1.0)
        UR = po.bc_ghost(bc.BC_FAR, UL, N_, FAR, GAMMA)
        rL, vL, pL = _prim(UL)
        rF, vF, pF = _prim(FAR)
        rR, vR, pR = _prim(UR)
        if case == "sup_out":
            np.testing.assert_allclose(UR, UL, rtol=0, atol=0)
        elif case == "sup_in":
            np.testing.assert_allclose(UR, FAR, rtol=0, atol=0)
        elif case == "sub_out":
            np.testing.assert_allclose([rR, *vR, pR], [rL, *vL, pF], rtol=1e-14, atol=1e-14)
        else:
            np.testing.assert_allclose([rR, *vR, pR], [rF, *vF, pL], rtol=1e-14, atol=1e-14)
    # BCOutP takes the same characteristic switch
    UL = _cons(1.0, 0.3 * N_, 1.0)
    np.testing.assert_array_equal(po.bc_ghost(bc.BC_OUTP, UL, N_, FAR), po.bc_ghost(bc.BC_FAR, UL, N_, FAR))


@pytest.mark.parametrize("t", [bc.BC_WALL_INVIS, bc.BC_SYM])
def test_slip_wall_reverses_normal_momentum(t):
    UL = _cons(1.1, [0.3, -0.2, 0.1], 0.8)
    UR = po.bc_ghost(t, UL, N_, None)
    assert UR[0] == UL[0] and UR[4] == UL[4]
    mL, mR = UL[1:4], UR[1:4]
    assert abs((mL + mR) @ N_) < 1e-15                       # normal momentum reversed
    t1 = np.array([-0.8, 0.6, 0.0])
    t2 = np.array([0.0, 0.0, 1.0])
    np.testing.assert_allclose([mR @ t1, mR @ t2], [mL @ t1, mL @ t2], atol=1e-15)


def test_wall_out_in_ghosts():
    UL = _cons(1.1, [0.3, -0.2, 0.1], 0.8)
    UR = po.bc_ghost(bc.BC_WALL, UL, N_, None)
    np.testing.assert_array_equal(UR, [UL[0], -UL[1], -UL[2], -UL[3], UL[4]])
    np.testing.assert_array_equal(po.bc_ghost(bc.BC_OUT, UL, N_, FAR), UL)
    np.testing.assert_array_equal(po.bc_ghost(bc.BC_IN, UL, N_, FAR), FAR)


# ---------------------------------------------------------------------------- VR with Dirichlet values
def _box(kind, jitter=0.12, seed=4):
    if kind == "2d":
        return mm.box_quad_grid(6, 5, 3.0, 2.5, x0=-1.0, y0=-1.0, zones=dict(xmin=0, xmax=0, ymin=0, ymax=0),
                                jitter=jitter, seed=seed)
    return mm.box_hex_grid(4, 4, 3, 2.0, 2.0, 1.5, x0=-1.0, y0=-1.0, z0=-0.5,
                           zones=dict(xmin=0, xmax=0, ymin=0, ymax=0, zmin=0, zmax=0), jitter=jitter, seed=seed)


def _dirichlet_values(o, f):
    pts = o.face_metrics()["pts"]
    return f(pts)[..., None]  # [nFace, G, 1]


FIELDS = {
    0: (lambda P: np.full(P.shape[:-1], 1.0), [1.0], 0),
    1: (lambda P: 0.5 + P[..., 0] + 2 * P[..., 1] - 0.5 * P[..., 2], [0.5, 1.0, 2.0, -0.5], 1),
    4: (lambda P: P[..., 0] ** 2 + P[..., 1] ** 2 + P[..., 2] ** 2, [1.0], 2),
    5: (lambda P: P[..., 0] ** 3 + P[..., 0] * P[..., 1] ** 2 + P[..., 2] ** 3, [1.0], 3),
}


def _run_vr(o, fid, nIter=300):
    """Jacobi VR sweeps with the exact field's values at the boundary face points, from a
    zero reconstruction; returns the L1 error of the reconstruction at cell quadrature points."""
    f, prm, _ = FIELDS[fid]
    o.set_vr_dirichlet(_dirichlet_values(o, f))   # 2-D points have z = 0
    u = o.cell_average(fid, prm, 1)
    rec = np.zeros((o.nCell, 1, o.M))
    for _ in range(nIter):
        new = o.vr_iter(rec, u)
        done = np.abs(new - rec).max() < 1e-15
        rec = new
        if done:
            break
    return o.rec_error_l1(fid, prm, u, rec)


@pytest.mark.parametrize("kind,order", [("2d", 1), ("2d", 2), ("2d", 3), ("3d", 1), ("3d", 2)])
def test_vr_dirichlet_constant_any_weights(kind, order):
    m = _box(kind)
    o = po.OracleVR(m, maxOrder=order, intOrder=max(order + 2, 3))
    assert _run_vr(o, 0) < 1e-12


@pytest.mark.parametrize("kind,order", [("2d", 1), ("2d", 2), ("2d", 3), ("3d", 1), ("3d", 2)])
def test_vr_dirichlet_polynomial_exactness(kind, order):
    m = _box(kind)
    zw = bc.euler_dir_weights([bc.BC_FAR], order)  # value-only boundary weights
    o = po.OracleVR(m, maxOrder=order, intOrder=max(order + 2, 3), zoneDirWeight=zw)
    for fid in (1, 4, 5):
        if FIELDS[fid][2] > order:
            continue
        err = _run_vr(o, fid, nIter=2000)
        assert err < 1e-9, (fid, err)


def test_vr_dirichlet_all_order_weights_are_not_exact_for_linear():
    """Sensitivity of the pin above: derivative terms on boundary faces without derivative
    data in the boundary term break exactness (so a missing or extra factor would too)."""
    m = _box("2d")
    o = po.OracleVR(m, maxOrder=2, intOrder=4)
    assert _run_vr(o, 1, nIter=2000) > 1e-6


# ---------------------------------------------------------------------------- EvaluateRHS with BCs
def _uniform(n, U):
    return np.tile(U, (n, 1))


@pytest.mark.parametrize("rs", ["Roe", "HLLEP", "HLLC", "Roe_M1"])
@pytest.mark.parametrize("order", [1, 2])
def test_free_stream_preserved_by_far_field_and_tangential_walls(rs, order):
    U = _cons(1.0, [0.5, 0.0, 0.0], 1.0 / GAMMA)
    # channel: far field at x ends, slip walls at y ends (flow tangential)
    m = mm.box_quad_grid(8, 5, 4.0, 2.0, zones=dict(xmin=0, xmax=0, ymin=1, ymax=1), jitter=0.0)
    types = [bc.BC_FAR, bc.BC_WALL_INVIS]
    o = po.OracleVR(m, maxOrder=order, intOrder=order + 2, zoneDirWeight=bc.euler_dir_weights(types, order))
    o.set_bc(types, np.stack([U, U]), GAMMA)
    u = _uniform(o.nCell, U)
    rec = o.vr_iter(np.zeros((o.nCell, 5, o.M)), u)
    assert np.abs(rec).max() < 1e-13          # the boundary values agree with the state: no slope
    es = o.euler_settings(GAMMA, RS[rs], 1 if rs == "Roe_M1" else 0)
    _, _, dlc, _ = o.eval_dt(es, u, 1.0)
    r = o.eval_rhs(es, u, rec, dlc)
    assert np.abs(r).max() < 1e-12


def _mirror_pair(jitter=0.08):
    """Half domain [0,4]x[0,2] with a symmetry plane at y = 0, and the full domain
    [0,4]x[-2,2] built as its exact mirror image (nodes mapped y -> -y)."""
    nx, ny = 6, 4

    def bend(P):  # smooth, y-odd node displacement: the mesh stays mirror symmetric
        Q = P.copy()
        Q[:, 0] += jitter * np.sin(1.3 * P[:, 0]) * np.cos(0.7 * P[:, 1]) * np.sin(0.9 * P[:, 1]) ** 2
        Q[:, 1] += jitter * np.sin(0.8 * P[:, 1]) * np.cos(1.1 * P[:, 0])
        return Q

    half = mm.box_quad_grid(nx, ny, 4.0, 2.0, zones=dict(xmin=0, xmax=0, ymin=1, ymax=0), node_map=bend)
    full = mm.box_quad_grid(nx, 2 * ny, 4.0, 4.0, y0=-2.0, zones=dict(xmin=0, xmax=0, ymin=0, ymax=0),
                            node_map=bend)
    upper = np.array([i + nx * (j + ny) for j in range(ny) for i in range(nx)])
    lower = np.array([i + nx * (ny - 1 - j) for j in range(ny) for i in range(nx)])
    return half, full, upper, lower


@pytest.mark.parametrize("rs", ["Roe", "HLLEP", "HLLC"])
@pytest.mark.parametrize("order", [1, 2])
def test_symmetry_plane_equals_mirrored_domain(rs, order):
    half, full, upper, lower = _mirror_pair()
    types_h, types_f = [bc.BC_FAR, bc.BC_SYM], [bc.BC_FAR]
    far = _cons(1.0, [0.4, 0.0, 0.0], 1.0 / GAMMA)  # no y velocity: the full domain is symmetric too
    oh = po.OracleVR(half, maxOrder=order, intOrder=order + 2, zoneDirWeight=bc.euler_dir_weights(types_h, order))
    of = po.OracleVR(full, maxOrder=order, intOrder=order + 2, zoneDirWeight=bc.euler_dir_weights(types_f, order))
    oh.set_bc(types_h, np.stack([far, far]), GAMMA)
    of.set_bc(types_f, far[None], GAMMA)
    rng = np.random.default_rng(5)
    n = oh.nCell
    uh = np.stack([_cons(rng.uniform(0.9, 1.1), [rng.uniform(-.2, .2), rng.uniform(-.2, .2), 0.0],
                         rng.uniform(0.6, 0.8)) for _ in range(n)])
    mirror = np.array([1, 1, -1, 1, 1.0])
    uf = np.zeros((of.nCell, 5))
    uf[upper] = uh
    uf[lower] = uh * mirror
    # reconstruction: random in the upper half, mirrored below (y-odd modes flip, rho v flips)
    Mh = oh.M
    # 2-D monomials in the reference order: x, y, x^2, xy, y^2, x^3, x^2y, xy^2, y^3
    ordy = {1: [0, 1], 2: [0, 1, 0, 1, 2], 3: [0, 1, 0, 1, 2, 0, 1, 2, 3]}[order]
    recH = rng.uniform(-0.02, 0.02, (n, 5, Mh))
    sgn = np.array([[mirror[v] * (-1) ** ordy[k] for k in range(Mh)] for v in range(5)])
    recF = np.zeros((of.nCell, 5, Mh))
    recF[upper] = recH
    recF[lower] = recH * sgn
    es = of.euler_settings(GAMMA, RS[rs])
    rh = oh.eval_rhs(es, uh, recH, None)
    rf = of.eval_rhs(es, uf, recF, None)
    scale = np.abs(rf[upper]).max()
    assert np.abs(rh - rf[upper]).max() <= 1e-12 * scale
    # the full-domain residual is itself mirror symmetric (a check on the construction)
    assert np.abs(rf[lower] - rf[upper] * mirror).max() <= 1e-12 * scale


def test_zone_force_of_uniform_pressure():
    """A uniform state moving parallel to a flat slip wall: the Riemann flux through the wall
    is p n, so the oracle's force on the wall zone is p times the wall length along its outward
    normal, (0, -p Lx, 0)."""
    Lx, p = 3.0, 0.9
    m = mm.box_quad_grid(6, 4, Lx, 2.0, zones=dict(xmin=0, xmax=0, ymin=1, ymax=0), jitter=0.0)
    types = [bc.BC_FAR, bc.BC_WALL_INVIS]
    U = _cons(1.1, [0.4, 0.0, 0.0], p)
    o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o.set_bc(types, np.stack([U, U]), 1.4)
    es = o.euler_settings(1.4, RS["Roe"])
    u = np.tile(U, (o.nCell, 1))
    f = o.zone_force(es, u, np.zeros((o.nCell, 5, o.M)), None, 1)
    np.testing.assert_allclose(f, [0.0, -p * Lx, 0.0], atol=1e-12)


# ------------------------------------------------------ isothermal wall and total-condition inflow
R_GAS = 0.8


def _prim(U, gamma=GAMMA):
    v = U[1:4] / U[0]
    return U[0], v, (gamma - 1) * (U[4] - 0.5 * U[0] * v @ v)


def test_isothermal_wall_ghost():
    """generateBV_ViscousWall with BCWallIsothermal (EvaluateDt.hxx:2241-2280): velocity
    reversed, the interior pressure, and the density that puts the ghost at the wall
    temperature (p = rho Rgas T_wall); at the interior temperature it is the adiabatic wall's."""
    UL = _cons(1.1, [0.3, -0.2, 0.1], 0.8)
    Tw = 1.3
    UR = po.bc_ghost(bc.BC_WALL_ISOTHERMAL, UL, N_, [Tw, 0, 0, 0, 0], GAMMA, R_GAS)
    rL, vL, pL = _prim(UL)
    rR, vR, pR = _prim(UR)
    assert abs(pR - pL) <= 1e-14 * pL
    np.testing.assert_allclose(vR, -vL, rtol=1e-14)
    assert abs(pR / (rR * R_GAS) - Tw) <= 1e-14 * Tw
    Tin = pL / (rL * R_GAS)
    np.testing.assert_allclose(po.bc_ghost(bc.BC_WALL_ISOTHERMAL, UL, N_, [Tin, 0, 0, 0, 0], GAMMA, R_GAS),
                               po.bc_ghost(bc.BC_WALL, UL, N_, None), rtol=1e-14, atol=1e-15)


def test_total_condition_inflow_ghost():
    """generateBV_TotalConditionInflow (EvaluateDt.hxx:2417-2455): the interior speed along the
    given direction, and the static state whose total temperature and total pressure are the
    given ones (isentropic: T0 = T + v^2/(2 Cp), p0 = p (T0/T)^(gamma/(gamma-1))); a speed above
    the limit sqrt(0.95 * 2 Cp T0) is clamped to it."""
    Cp = GAMMA / (GAMMA - 1) * R_GAS
    p0, T0, d = 1.2, 0.9, np.array([2.0, 1.0, -0.5])
    for speed in (0.3, 1.5, 5.0):
        UL = _cons(1.1, speed * np.array([0.6, -0.8, 0.0]), 0.8)
        UR = po.bc_ghost(bc.BC_IN_PSTS, UL, N_, [p0, T0, *d], GAMMA, R_GAS)
        r, v, p = _prim(UR)
        vmax = np.sqrt(0.95 * 2 * Cp * T0)
        s = min(speed, vmax)
        np.testing.assert_allclose(v, s * d / np.linalg.norm(d), rtol=1e-13)
        T = p / (r * R_GAS)
        assert abs(T + s * s / (2 * Cp) - T0) <= 1e-13 * T0
        assert abs(p * (T0 / T) ** (GAMMA / (GAMMA - 1)) - p0) <= 1e-12 * p0


def test_isothermal_wall_conducts_heat():
    """A gas at rest next to an isothermal wall: no residual when the wall is at the gas
    temperature; a hotter wall heats the cells along it (positive energy residual there,
    momentum along the wall zero), a colder one cools them, the adiabatic wall does neither."""
    m = mm.box_quad_grid(6, 5, 3.0, 2.5, zones=dict(xmin=0, xmax=1, ymin=1, ymax=1), jitter=0.0, seed=1)
    types = [bc.BC_WALL_ISOTHERMAL, bc.BC_SYM]
    o = po.OracleVR(m, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
    rho, p = 1.0, 1.0 / GAMMA
    Tgas = p / (rho * R_GAS)
    u = np.tile(_cons(rho, [0.0, 0.0, 0.0], p), (o.nCell, 1))
    es = o.euler_settings(GAMMA, RS["Roe"], viscous=1, muModel=0, Rgas=R_GAS, muGas=0.05, prGas=0.72)
    wallCells = np.unique(m.face2cell[(m.face2cell[:, 1] < 0) & (m.faceZone == 0), 0])
    res = {}
    for Tw in (Tgas, 1.5 * Tgas, 0.5 * Tgas):
        o.set_bc(types, np.tile([Tw, 0, 0, 0, 0], (2, 1)), GAMMA, R_GAS)
        _, _, dlc, _ = o.eval_dt(es, u, 1.0, 1e100, True)
        res[Tw] = o.eval_rhs(es, u, None, dlc)
    assert np.abs(res[Tgas]).max() <= 1e-13
    hot, cold = res[1.5 * Tgas], res[0.5 * Tgas]
    assert np.all(hot[wallCells, 4] > 0) and np.all(cold[wallCells, 4] < 0)
    o.set_bc([bc.BC_WALL, bc.BC_SYM], None, GAMMA, R_GAS)
    assert np.abs(o.eval_rhs(es, u, None, dlc)).max() <= 1e-13
