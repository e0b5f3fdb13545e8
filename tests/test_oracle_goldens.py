"""Pins the CPU oracle against the reference's own golden values and against
closed-form / brute-force properties (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import mesh as mm
from helpers import RS, random_state, rel_l2

G = 1.4


def p2c(r, u, v, w, p):
    return np.array([r, r * u, r * v, r * w, p / (G - 1) + 0.5 * r * (u * u + v * v + w * w)])


def exact_flux(U, n):
    v = U[1:4] / U[0]
    p = (G - 1) * (U[4] - 0.5 * U[0] * v @ v)
    vn = v @ n
    return np.concatenate([[U[0] * vn], U[1:4] * vn + p * n, [(U[4] + p) * vn]])


# ---------------------------------------------------------------- Riemann solvers
GOLDEN_UL = p2c(1.225, 100.0, -50.0, 25.0, 101325.0)
GOLDEN_UR = p2c(0.8, 200.0, 0.0, 0.0, 80000.0)
GOLDEN_N = np.array([0.6, 0.8, 0.0]) / np.linalg.norm([0.6, 0.8, 0.0])


@pytest.mark.parametrize("rs,golden", [
    # test/cpp/Euler/test_RiemannSolvers.cpp:257-267 (dLambda = fixScale = incFScale = 0)
    ("Roe", [8.1162345145e+01, 5.8813861917e+04, 5.8759839011e+04, 5.9539454795e+02, 2.7251925995e+07]),
    ("HLLC", [9.3486183606e+01, 5.5647501399e+04, 5.7057534871e+04, 2.3371545902e+03, 2.5454872499e+07]),
    ("HLLEP", [9.2861523292e+01, 5.1385111929e+04, 6.2875774679e+04, 5.7376094955e+03, 2.6719441912e+07]),
])
def test_riemann_golden(rs, golden):
    F, _ = po.riemann(RS[rs], GOLDEN_UL, GOLDEN_UR, GOLDEN_N, G, dLambda=0.0, fixScale=0.0, incFScale=0.0)
    np.testing.assert_allclose(F, golden, rtol=1e-8)


@pytest.mark.parametrize("rs", list(RS))
def test_riemann_consistency_and_symmetry(rs):
    rng = np.random.default_rng(5)
    for _ in range(20):
        U = p2c(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        F, _ = po.riemann(RS[rs], U, U, n, G, dLambda=0.1)
        np.testing.assert_allclose(F, exact_flux(U, n), rtol=1e-12, atol=1e-12)
        # mirror symmetry F(UL, UR, n) = -F(UR, UL, -n)
        U2 = p2c(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2))
        F1, _ = po.riemann(RS[rs], U, U2, n, G, dLambda=0.1)
        F2, _ = po.riemann(RS[rs], U2, U, -n, G, dLambda=0.1)
        np.testing.assert_allclose(F1, -F2, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("rs", ["Roe", "HLLEP", "HLLC"])
def test_riemann_supersonic_is_upwind(rs):
    # both states supersonic to the right: every upwind solver returns F(UL)
    UL = p2c(1.0, 3.0, 0.1, 0.0, 1.0)
    UR = p2c(0.9, 3.1, 0.0, 0.2, 1.1)
    n = np.array([1.0, 0.0, 0.0])
    F, _ = po.riemann(RS[rs], UL, UR, n, G)
    np.testing.assert_allclose(F, exact_flux(UL, n), rtol=1e-12)


def test_roe_harten_yee_closed_form():
    # Roe (eigScheme 0) equals 0.5(FL+FR) - 0.5 R |Lambda|_fixed L dU; build R, L from the
    # textbook Euler eigensystem and check against the solver at a sonic-point state.
    UL = p2c(1.0, 0.95, 0.0, 0.0, 1.0)
    UR = p2c(0.8, 1.25, 0.1, 0.0, 0.7)
    n = np.array([1.0, 0.0, 0.0])
    F, lam = po.riemann(RS["Roe"], UL, UR, n, G, fixScale=1.0, incFScale=1.0)
    sL, sR = np.sqrt(UL[0]), np.sqrt(UR[0])
    vL, vR = UL[1:4] / UL[0], UR[1:4] / UR[0]
    pL = (G - 1) * (UL[4] - 0.5 * UL[0] * vL @ vL)
    pR = (G - 1) * (UR[4] - 0.5 * UR[0] * vR @ vR)
    HL, HR = (UL[4] + pL) / UL[0], (UR[4] + pR) / UR[0]
    v = (sL * vL + sR * vR) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    a = np.sqrt((G - 1) * (H - 0.5 * v @ v))
    u = v[0]
    R = np.array([[1, 1, 0, 0, 1], [u - a, u, 0, 0, u + a], [v[1], v[1], 1, 0, v[1]], [v[2], v[2], 0, 1, v[2]],
                  [H - u * a, 0.5 * v @ v, v[1], v[2], H + u * a]])
    aL = np.sqrt(G * pL / UL[0])
    aR = np.sqrt(G * pR / UR[0])

    def hy(lamv, l, lL, lR):
        th = max(0.0, l - lL, lR - l) * 0.05
        return (lamv ** 2 / th + th) / 2 if lamv < th else lamv
    lams = [hy(abs(u - a), u - a, vL[0] - aL, vR[0] - aR), hy(abs(u), u, vL[0], vR[0]),
            hy(abs(u), u, vL[0], vR[0]), hy(abs(u), u, vL[0], vR[0]), hy(abs(u + a), u + a, vL[0] + aL, vR[0] + aR)]
    ref = 0.5 * (exact_flux(UL, n) + exact_flux(UR, n)) - 0.5 * R @ np.diag(lams) @ np.linalg.solve(R, UR - UL)
    np.testing.assert_allclose(F, ref, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(lam, [lams[0], lams[1], lams[4]], rtol=1e-12)


# ---------------------------------------------------------------- VR goldens
def _converge(o, fid, prm, maxit=200):
    u = o.cell_average(fid, prm, 1)
    rec = np.zeros((o.nCell, 1, o.M))
    for _ in range(maxit):
        nw = o.vr_iter(rec, u)
        inc = np.sqrt(((nw - rec) ** 2).sum() / o.nCell)
        rec = nw
        if inc < 1e-14:
            break
    return o.rec_error_l1(fid, prm, u, rec)


HQM = dict(dirWeightScheme=1, geomWeightScheme=1, geomWeightBias=1.0, geomWeightPower=0.5)


@pytest.mark.parametrize("n,order,hqm,fid,golden", [
    # test/cpp/CFV/test_Reconstruction.cpp:531-545, IV10 quads (uniform 10x10 on [0,10]^2 and bisections)
    (10, 1, False, 2, 4.6604402914e-02), (20, 1, False, 2, 9.2961629825e-03), (40, 1, False, 2, 1.5347312301e-03),
    (10, 1, True, 2, 4.6604402914e-02), (20, 1, True, 2, 9.2961629825e-03),
    (10, 2, True, 2, 3.0528143687e-03), (20, 2, True, 2, 2.3057099673e-04), (40, 2, True, 2, 2.4367733525e-05),
    (10, 3, True, 2, 1.9105219870e-03), (20, 3, True, 2, 4.6701352192e-05), (40, 3, True, 2, 1.4890868814e-06),
    (10, 3, False, 2, 1.8840503911e-03), (20, 3, False, 2, 2.4995731251e-05), (40, 3, False, 2, 8.3670061853e-07),
    (10, 3, True, 3, 5.5612138851e-04), (20, 3, True, 3, 1.6082305163e-05), (40, 3, True, 3, 6.6188225747e-07),
])
def test_vr_golden_2d(n, order, hqm, fid, golden):
    m = mm.periodic_quad_grid(n, n, 10.0, 10.0)
    o = po.OracleVR(m, maxOrder=order, intOrder=max(order + 2, 5), **(HQM if hqm else {}))
    err = _converge(o, fid, [10.0, 2])
    assert abs(err - golden) <= 1e-6 * golden


@pytest.mark.parametrize("order,fid,golden", [
    # test/cpp/CFV/test_Reconstruction3D.cpp:393-409, Uniform32_3D_Periodic on [0,1]^3
    (1, 2, 3.2475021825e-03), (2, 2, 7.4823241394e-05), (1, 3, 2.4261881197e-03), (2, 3, 3.7014576777e-05),
])
def test_vr_golden_3d(order, fid, golden):
    m = mm.periodic_hex_grid(32, 32, 32, 1.0, 1.0, 1.0)
    o = po.OracleVR(m, maxOrder=order, intOrder=5, **HQM)
    err = _converge(o, fid, [1.0, 3])
    assert abs(err - golden) <= 1e-6 * golden


@pytest.mark.parametrize("n,golden", [
    # test/cpp/CFV/test_Reconstruction.cpp:533: Gauss-Green (subs2ndOrder 1, distance-weighted
    # face values), IV10 quads and two bisections, sin(kx)cos(ky)
    (10, 1.5599270188e-02), (20, 3.4233789068e-03), (40, 7.8943623278e-04)])
def test_gauss_green_golden_2d(n, golden):
    """The reference measures u_i + grad_i.(x - bary_i); the substitute's linear modes are
    grad * scale on the zero-mean basis (x - bary) / scale, the same polynomial."""
    m = mm.periodic_quad_grid(n, n, 10.0, 10.0)
    o = po.OracleVR(m, maxOrder=1, intOrder=5, subs2ndOrder=1)
    u = o.cell_average(2, [10.0, 2], 1)
    err = o.rec_error_l1(2, [10.0, 2], u, o.vr_2nd(u))
    assert abs(err - golden) <= 1e-6 * golden


@pytest.mark.parametrize("fid,golden", [
    # test/cpp/CFV/test_Reconstruction3D.cpp:395-397,401-403: Gauss-Green on Uniform32_3D_Periodic
    (2, 1.5098818118e-03), (3, 1.4931684237e-03)])
def test_gauss_green_golden_3d(fid, golden):
    m = mm.periodic_hex_grid(32, 32, 32, 1.0, 1.0, 1.0)
    o = po.OracleVR(m, maxOrder=1, intOrder=5, subs2ndOrder=1)
    u = o.cell_average(fid, [1.0, 3], 1)
    err = o.rec_error_l1(fid, [1.0, 3], u, o.vr_2nd(u))
    assert abs(err - golden) <= 1e-6 * golden


@pytest.mark.parametrize("order", [1, 2, 3])
def test_vr_constant_and_linear_exactness(order):
    m = mm.periodic_quad_grid(6, 5, 6.0, 5.0, jitter=0.2, seed=3)
    o = po.OracleVR(m, maxOrder=order, intOrder=5)
    u = o.cell_average(0, [2.5], 1)
    rec = np.zeros((o.nCell, 1, o.M))
    for _ in range(50):
        rec = o.vr_iter(rec, u)
    assert np.abs(rec).max() < 1e-13
    assert o.rec_error_l1(0, [2.5], u, rec) < 1e-13


def test_vr_fixed_point_is_dense_solution():
    """The Jacobi iteration converges to the solution of the dense block system
    (I - sum A^-1 B) x = sum A^-1 b du, assembled by brute force."""
    m = mm.periodic_quad_grid(4, 3, 4.0, 3.0, jitter=0.2, seed=1)
    o = po.OracleVR(m, maxOrder=2, intOrder=4)
    N, M = o.nCell, o.M
    A, b = o.vr_coeffs()
    u = np.random.default_rng(0).uniform(0, 1, (N, 1))
    K = np.eye(N * M)
    rhs = np.zeros(N * M)
    for i in range(N):
        for k, f in enumerate(m.cell2face[4 * i:4 * i + 4]):
            j = m.face2cell[f, 1] if m.face2cell[f, 0] == i else m.face2cell[f, 0]
            K[i * M:(i + 1) * M, j * M:(j + 1) * M] -= A[i, 1 + k]
            rhs[i * M:(i + 1) * M] += b[i, k] * (u[j, 0] - u[i, 0])
    x = np.linalg.solve(K, rhs).reshape(N, 1, M)
    rec = np.zeros((N, 1, M))
    for _ in range(400):
        rec = o.vr_iter(rec, u)
    assert rel_l2(rec, x) < 1e-10
    # and the recordInc-free Jacobian action vanishes on the homogeneous part
    np.testing.assert_allclose(o.vr_iter_diff(x) .ravel(), (K @ x.ravel()).reshape(N, 1, M).ravel(), atol=1e-13)


# ---------------------------------------------------------------- Euler pipeline golden
def test_iv_pipeline_golden():
    """test/cpp/Euler/test_EulerEvaluator.cpp:231-259: IV (chi 10.0828) on 10x10 periodic quads,
    p1 / intOrder 3, Roe_M1 batched, local dt at CFL 100, first-order RHS, LU-SGS init and
    block-Jacobi increment."""
    m = mm.periodic_quad_grid(10, 10, 10.0, 10.0)
    o = po.OracleVR(m, maxOrder=1, intOrder=3)
    u = o.cell_average(10, [10.0828, 1.4, 1.0, 1.0], 5)
    es = o.euler_settings(1.4, RS["Roe_M1"], 1)
    dt, lam, dlc, _ = o.eval_dt(es, u, 100.0, 1e100, True)
    rhs = o.eval_rhs(es, u, None, dlc)
    vol = o.cell_metrics()["vol"]
    res = (np.abs(rhs) * vol[:, None]).sum(0)
    assert abs(res[0] - 1.0524823780e+01) <= 1e-6 * 1.0524823780e+01
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    z = np.zeros_like(u)
    x, _ = o.update_sgs(es, 1.0, rhs, u, z, z, jd, lam, True, False, False)
    inc = np.linalg.norm(np.sqrt((x ** 2).sum(0)))
    assert abs(inc - 3.1557730986e+00) <= 1e-6 * 3.1557730986e+00


def box_state(o):
    """cases/euler3D/euler3D_config_Box.json: far field (1, 0, 0, 0, 2.5), and
    (0.5, 1, 0, 0, 4) in cells whose barycentre lies strictly inside (0.25, 0.75)^3
    (boxInitializers, src/Euler/EulerEvaluator.hxx:1204-1218)."""
    bary = o.cell_metrics()["bary"] if hasattr(o, "nCell") else o
    u = np.tile([1.0, 0.0, 0.0, 0.0, 2.5], (len(bary), 1))
    u[np.all((bary > 0.25) & (bary < 0.75), 1)] = [0.5, 1.0, 0.0, 0.0, 4.0]
    return u


def test_box_pipeline_golden():
    """test/cpp/Euler/test_EulerEvaluator.cpp:297-330: 3-D box (NS_3D) on the 32^3 periodic
    unit cube, p1 / intOrder 3, HLLC, local dt at CFL 10 (euler3D_config_Box.json), first-order
    EvaluateRHS, LUSGSMatrixInit (dt 1e100) and the block-Jacobi increment."""
    m = mm.periodic_hex_grid(32, 32, 32, 1.0, 1.0, 1.0)
    o = po.OracleVR(m, maxOrder=1, intOrder=3)
    u = box_state(o)
    es = o.euler_settings(1.4, RS["HLLC"])
    dt, lam, dlc, _ = o.eval_dt(es, u, 10.0, 1e100, True)
    rhs = o.eval_rhs(es, u, None, dlc)
    vol = o.cell_metrics()["vol"]
    res = (np.abs(rhs) * vol[:, None]).sum(0)
    assert abs(res[0] - 5.8855408029e-01) <= 1e-6 * 5.8855408029e-01
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    z = np.zeros_like(u)
    x, _ = o.update_sgs(es, 1.0, rhs, u, z, z, jd, lam, True, False, False)
    inc = np.linalg.norm(np.sqrt((x ** 2).sum(0)))
    assert abs(inc - 3.4532977412e+01) <= 1e-6 * 3.4532977412e+01


# ---------------------------------------------------------------- RHS / SGS properties
@pytest.mark.parametrize("rs", ["Roe", "HLLEP", "HLLC", "Roe_M1"])
def test_rhs_conservation_and_freestream(rs):
    m = mm.periodic_hex_grid(4, 3, 3, 1.0, 1.0, 1.0, jitter=0.15, seed=2)
    o = po.OracleVR(m, maxOrder=2, intOrder=5)
    es = o.euler_settings(1.4, RS[rs])
    N = o.nCell
    u = random_state(N, 9)
    rec = np.random.default_rng(1).uniform(-1e-2, 1e-2, (N, 5, o.M))
    _, _, dlc, _ = o.eval_dt(es, u, 1.0)
    rhs = o.eval_rhs(es, u, rec, dlc)
    vol = o.cell_metrics()["vol"]
    scale = np.abs(rhs * vol[:, None]).sum()
    assert np.abs((rhs * vol[:, None]).sum(0)).max() <= 1e-13 * scale
    uc = np.tile(random_state(1, 3), (N, 1))
    r0 = o.eval_rhs(es, uc, np.zeros_like(rec), np.zeros(N))
    assert np.abs(r0).max() < 1e-12


def test_flux_jacobian_increment_matches_fd():
    """The LU-SGS off-diagonal term dF = A(U)dU.n - lambda dU: with lambda = 0 it must equal the
    directional derivative of the physical normal flux (central differences)."""
    m = mm.periodic_quad_grid(3, 3, 3.0, 3.0)
    o = po.OracleVR(m, maxOrder=1, intOrder=1)
    es = o.euler_settings(1.4, 1)
    N = o.nCell
    u = random_state(N, 4)
    dU = np.random.default_rng(2).normal(size=(N, 5))
    # single nonzero neighbour increment: cell 4's backward GS sweep sees only its neighbours
    rhs = np.zeros((N, 5))
    jd = np.ones(N)
    lam = np.zeros((o.nFace, 4))
    out, _ = o.update_sgs(es, 1.0, rhs, u, dU, np.zeros_like(u), jd, lam, True, False, False)
    fm = o.face_metrics()
    vol = o.cell_metrics()["vol"]
    i = 4
    ref = np.zeros(5)
    for f in m.cell2face[4 * i:4 * i + 4]:
        L, R = m.face2cell[f]
        j = R if L == i else L
        n = fm["meanNorm"][f] * (1 if L == i else -1)
        eps = 1e-6
        d = (exact_flux(u[j] + eps * dU[j], n) - exact_flux(u[j] - eps * dU[j], n)) / (2 * eps)
        ref -= 0.5 * fm["area"][f] / vol[i] * d
    np.testing.assert_allclose(out[i], ref, rtol=1e-7, atol=1e-9)


def test_sgs_symmetric_sweep_matches_dense_gauss_seidel():
    """Forward GS sweep from zero == forward substitution on the dense (D + L) system."""
    m = mm.periodic_quad_grid(4, 4, 4.0, 4.0, jitter=0.1, seed=5)
    o = po.OracleVR(m, maxOrder=1, intOrder=1)
    es = o.euler_settings(1.4, 1)
    N = o.nCell
    u = random_state(N, 6)
    dt, lam, _, _ = o.eval_dt(es, u, 2.0)
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    rhs = np.random.default_rng(3).uniform(-1, 1, (N, 5))
    order = np.random.default_rng(4).permutation(N).astype(np.int64)
    # dense operator: row block i = D_i I + sum_j c_ij J_j with J_j = dF/du_j (linear in du)
    Jop = np.zeros((N * 5, N * 5))
    for j in range(N):
        for k in range(5):
            e = np.zeros((N, 5))
            e[j, k] = 1.0
            y, _ = o.update_sgs(es, 1.0, np.zeros((N, 5)), u, e, np.zeros((N, 5)), np.ones(N), lam, True, False, False)
            Jop[:, j * 5 + k] = -y.ravel()      # y_i = -sum_j c_ij dF(e_j) / 1
    pos = np.empty(N, int)
    pos[order] = np.arange(N)
    Lmat = np.zeros_like(Jop)
    for i in range(N):
        for j in range(N):
            if pos[j] < pos[i]:
                Lmat[i * 5:(i + 1) * 5, j * 5:(j + 1) * 5] = Jop[i * 5:(i + 1) * 5, j * 5:(j + 1) * 5]
    D = np.kron(np.diag(jd), np.eye(5))
    ref = np.linalg.solve(D + Lmat, rhs.ravel()).reshape(N, 5)
    z = np.zeros((N, 5))
    x, _ = o.update_sgs(es, 1.0, rhs, u, z, z, jd, lam, True, True, True, order)
    assert rel_l2(x, ref) < 1e-13


def test_hllep_stationary_contact_closed_form():
    """HLLE+ (Gas.hpp:601-714) across a stationary contact (u = 0, equal p, rho_L != rho_R):
    the physical fluxes agree and the density jump is the contact wave alone
    (alpha_1 = d rho), so F = (0, p n, 0) + S+ S- / (S+ - S-) (1 - delta) d rho e_0 with
    S- = min(-a_Roe, -a_L), S+ = max(a_Roe, a_R) and delta = a_Roe / (a_Roe +
    max(|u_Roe|, dLambda fixScale)): pins the role of dLambda and fixScale."""
    g = 1.4
    rhoL, rhoR, p = 1.0, 0.6, 0.9
    n = np.array([0.36, -0.48, 0.8])
    UL = np.array([rhoL, 0, 0, 0, p / (g - 1)])
    UR = np.array([rhoR, 0, 0, 0, p / (g - 1)])
    HL, HR = g * p / ((g - 1) * rhoL), g * p / ((g - 1) * rhoR)
    HRoe = (np.sqrt(rhoL) * HL + np.sqrt(rhoR) * HR) / (np.sqrt(rhoL) + np.sqrt(rhoR))
    aRoe = np.sqrt((g - 1) * HRoe)
    Sm, Sp = min(-aRoe, -np.sqrt(g * p / rhoL)), max(aRoe, np.sqrt(g * p / rhoR))
    seen = set()
    for dLambda, fixScale in ((0.0, 1.0), (0.3, 1.0), (2.0, 1.0), (2.0, 0.5)):
        F, lam = po.riemann(RS["HLLEP"], UL, UR, n, g, dLambda=dLambda, fixScale=fixScale)
        delta = aRoe / (aRoe + max(0.0, dLambda * fixScale))
        expect = np.concatenate([[Sp * Sm / (Sp - Sm) * (1 - delta) * (rhoR - rhoL)], p * n, [0.0]])
        assert np.abs(F - expect).max() <= 1e-14, (dLambda, fixScale, F, expect)
        np.testing.assert_allclose(lam, [aRoe, 0.0, aRoe], rtol=1e-14, atol=1e-15)
        seen.add(round(F[0], 12))
    assert len(seen) == 4   # each (dLambda, fixScale) changes the contact dissipation
