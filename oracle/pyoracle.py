"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DNDSR_ORACLE_LIB selects another build of the same sources (the sanitizer build of
# oracle/Makefile's `sanitize` target)
_LIB = os.environ.get("DNDSR_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("nNode", C.c_int64), ("coords", dp), ("nCell", C.c_int64),
                ("cellType", i32p), ("cell2nodePtr", i64p), ("cell2node", i64p),
                ("cell2facePtr", i64p), ("cell2face", i64p), ("nFace", C.c_int64),
                ("faceType", i32p), ("face2nodePtr", i64p), ("face2node", i64p),
                ("face2cell", i64p), ("faceShift", dp), ("faceZone", i32p), ("nBCZone", C.c_int32)]


class VRSettings(C.Structure):
    _fields_ = [("maxOrder", C.c_int), ("intOrder", C.c_int), ("dirWeightScheme", C.c_int),
                ("geomWeightScheme", C.c_int), ("geomWeightBias", C.c_double),
                ("geomWeightPower", C.c_double), ("scaleMultiplier", C.c_double),
                ("jacobiRelax", C.c_double), ("nIterCellSmoothScale", C.c_int), ("bcWeight", C.c_double),
                ("zoneDirWeight", dp), ("SORInstead", C.c_int), ("subs2ndOrder", C.c_int),
                ("subs2ndOrderGGScheme", C.c_int)]


class EulerSettings(C.Structure):
    _fields_ = [("gamma", C.c_double), ("rsType", C.c_int), ("rsMeanValueEig", C.c_int),
                ("rsFixScale", C.c_double), ("rsIncFScale", C.c_double), ("viscous", C.c_int),
                ("muModel", C.c_int), ("Rgas", C.c_double), ("muGas", C.c_double), ("prGas", C.c_double),
                ("TRef", C.c_double), ("CSutherland", C.c_double), ("useRoeJacobian", C.c_int)]


class ImplicitSettings(C.Structure):
    _fields_ = [("dt", C.c_double), ("CFL", C.c_double), ("nRec", C.c_int), ("sgsIter", C.c_int),
                ("gsUpdate", C.c_int), ("maxIter", C.c_int), ("minIter", C.c_int), ("tol", C.c_double),
                ("fp32Sweeps", C.c_int), ("incrementRelax", C.c_double), ("refRho", C.c_double),
                ("refP", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.orc_vr_create.restype = C.c_void_p
        _lib.orc_vr_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(VRSettings)]
        _lib.orc_vr_destroy.argtypes = [C.c_void_p]
        _lib.orc_last_error.restype = C.c_char_p
        _lib.orc_rec_error_l1.restype = C.c_double
        for name in ["orc_vr_info", "orc_get_cell_metrics", "orc_get_face_metrics", "orc_get_cell_quad",
                     "orc_get_vr_coeffs", "orc_cell_average", "orc_vr_iter", "orc_vr_sor", "orc_vr_2nd", "orc_vr_iter_diff",
                     "orc_eval_dt", "orc_eval_rhs", "orc_eval_rhs_lam", "orc_zone_force", "orc_lusgs_init", "orc_update_sgs",
                     "orc_lusgs_init_block", "orc_update_sgs_block", "orc_update_sgs_transpose", "orc_ssprk3", "orc_implicit_euler_step", "orc_implicit_bdf_step", "orc_lusgs_matvec",
                     "orc_rec_error_l1", "orc_riemann", "orc_set_bc", "orc_set_vr_dirichlet", "orc_bc_ghost",
                     "orc_cell_rhs_alpha", "orc_compress_inc", "orc_add_increment_pp", "orc_model_rhs"]:
            getattr(_lib, name).argtypes = None
    return _lib


def set_threads(n, rhsGather=False):
    lib().orc_set_threads(C.c_int(n), C.c_int(int(rhsGather)))


def P(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def riemann(rsType, UL, UR, n, gamma=1.4, ULm=None, URm=None, dLambda=0.0, fixScale=1.0, incFScale=1.0):
    UL = np.ascontiguousarray(UL, np.float64)
    UR = np.ascontiguousarray(UR, np.float64)
    ULm = UL if ULm is None else np.ascontiguousarray(ULm, np.float64)
    URm = UR if URm is None else np.ascontiguousarray(URm, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    F = np.zeros(5)
    lam = np.zeros(3)
    _chk(lib().orc_riemann(C.c_int(rsType), C.c_int(3), P(UL), P(UR), P(ULm), P(URm), P(n), C.c_double(gamma),
                           C.c_double(dLambda), C.c_double(fixScale), C.c_double(incFScale), P(F), P(lam)))
    return F, lam


def cell_rhs_alpha(u, res, gamma=1.4, relax=0.9, rhoEps=1e-11, pEps=1e-11, compress=2):
    """EvaluateCellRHSAlpha of one cell: (alpha, limited); raises where the reference asserts."""
    a, lim = C.c_double(), C.c_int()
    rc = lib().orc_cell_rhs_alpha(P(np.ascontiguousarray(u, np.float64)), P(np.ascontiguousarray(res, np.float64)),
                                  C.c_double(gamma), C.c_double(relax), C.c_double(rhoEps), C.c_double(pEps),
                                  C.c_int(compress), C.byref(a), C.byref(lim))
    if rc:
        raise RuntimeError("oracle: EvaluateCellRHSAlpha assertion")
    return a.value, bool(lim.value)


def compress_inc(u, inc, gamma=1.4, rhoEps=1e-11, pEps=1e-11):
    """CompressInc of one cell's increment; raises where the reference asserts."""
    out = np.zeros(5)
    rc = lib().orc_compress_inc(P(np.ascontiguousarray(u, np.float64)), P(np.ascontiguousarray(inc, np.float64)),
                                C.c_double(gamma), C.c_double(rhoEps), C.c_double(pEps), P(out))
    if rc:
        raise RuntimeError("oracle: CompressInc assertion")
    return out


def add_increment_pp(u, du, gamma=1.4, relax=0.9, rhoEps=1e-11, pEps=1e-11):
    """fincrement over all cells (rows [N, 5]): returns (u_new, nLimited, alphaMin)."""
    u = np.array(u, dtype=np.float64, copy=True)
    du = np.ascontiguousarray(du, np.float64)
    am, bad = C.c_double(), C.c_int()
    f = lib().orc_add_increment_pp
    f.restype = C.c_int64
    n = f(C.c_int64(u.shape[0]), C.c_double(gamma), C.c_double(relax), C.c_double(rhoEps), C.c_double(pEps), P(u),
          P(du), C.byref(am), C.byref(bad))
    if bad.value:
        raise RuntimeError("oracle: positivity-preserving increment failed")
    return u, int(n), am.value


def bc_ghost(bcType, UL, n, zoneValue=None, gamma=1.4, Rgas=1.0):
    """Ghost state of one boundary point (generateBoundaryValue)."""
    UL = np.ascontiguousarray(UL, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    zv = np.zeros(5) if zoneValue is None else np.ascontiguousarray(zoneValue, np.float64)
    UR = np.zeros(5)
    _chk(lib().orc_bc_ghost(C.c_int(bcType), P(UL), P(n), P(zv), C.c_double(gamma), C.c_double(Rgas), P(UR)))
    return UR


class OracleVR:
    """VariationalReconstruction + EulerEvaluator restated on the CPU."""

    def __init__(self, mesh, maxOrder=1, intOrder=3, dirWeightScheme=0, geomWeightScheme=0,
                 geomWeightBias=0.0, geomWeightPower=0.5, jacobiRelax=1.0, nIterCellSmoothScale=15,
                 bcWeight=1.0, zoneDirWeight=None, SORInstead=False, subs2ndOrder=0, subs2ndOrderGGScheme=1):
        L = lib()
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(x) for x in (
            mesh.coords, mesh.cellType, mesh.cell2nodePtr, mesh.cell2node, mesh.cell2facePtr, mesh.cell2face,
            mesh.faceType, mesh.face2nodePtr, mesh.face2node, mesh.face2cell, mesh.faceShift)]
        k = self._keep
        fz = getattr(mesh, "faceZone", None)
        nz = int(getattr(mesh, "nBCZone", 0))
        fz = None if fz is None else np.ascontiguousarray(fz, np.int32)
        zw = None if zoneDirWeight is None else np.ascontiguousarray(zoneDirWeight, np.float64)
        self._keep += [fz, zw]
        md = MeshDesc(mesh.dim, mesh.nNode, k[0].ctypes.data_as(dp), mesh.nCell, k[1].ctypes.data_as(i32p),
                      k[2].ctypes.data_as(i64p), k[3].ctypes.data_as(i64p), k[4].ctypes.data_as(i64p),
                      k[5].ctypes.data_as(i64p), mesh.nFace, k[6].ctypes.data_as(i32p), k[7].ctypes.data_as(i64p),
                      k[8].ctypes.data_as(i64p), k[9].ctypes.data_as(i64p), k[10].ctypes.data_as(dp),
                      None if fz is None else fz.ctypes.data_as(i32p), nz)
        s = VRSettings(maxOrder, intOrder, dirWeightScheme, geomWeightScheme, geomWeightBias, geomWeightPower,
                       1.0, jacobiRelax, nIterCellSmoothScale, bcWeight,
                       None if zw is None else zw.ctypes.data_as(dp), int(bool(SORInstead)), int(subs2ndOrder),
                       int(subs2ndOrderGGScheme))
        self.settings = s
        h = L.orc_vr_create(C.byref(md), C.byref(s))
        if not h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        self.h = C.c_void_p(h)
        nC, nF = C.c_int64(), C.c_int64()
        M, mfc, gf, gc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        gv = C.c_double()
        L.orc_vr_info(self.h, C.byref(nC), C.byref(nF), C.byref(M), C.byref(mfc), C.byref(gf), C.byref(gc),
                      C.byref(gv))
        self.nCell, self.nFace, self.M, self.maxFC = nC.value, nF.value, M.value, mfc.value
        self.nGFace, self.nGCell, self.globalVol = gf.value, gc.value, gv.value
        self.maxOrder = maxOrder

    def __del__(self):
        try:
            lib().orc_vr_destroy(self.h)
        except Exception:
            pass

    # ---- metrics
    def cell_metrics(self):
        N, M = self.nCell, self.M
        vol, sc, sm = np.zeros(N), np.zeros(N), np.zeros(N)
        bary, cent, mom = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, M))
        _chk(lib().orc_get_cell_metrics(self.h, P(vol), P(bary), P(cent), P(sc), P(sm), P(mom)))
        return dict(vol=vol, bary=bary, cent=cent, scale=sc, smooth=sm, moment=mom)

    def face_metrics(self):
        F, G = self.nFace, self.nGFace
        area, fl = np.zeros(F), np.zeros(F)
        mn, pts, nrm = np.zeros((F, 3)), np.zeros((F, G, 3)), np.zeros((F, G, 3))
        wj, fw = np.zeros((F, G)), np.zeros((F, self.maxOrder + 1))
        _chk(lib().orc_get_face_metrics(self.h, P(area), P(mn), P(pts), P(nrm), P(wj), P(fl), P(fw)))
        return dict(area=area, meanNorm=mn, pts=pts, norms=nrm, wJDet=wj, faceL=fl, faceWeight=fw)

    def cell_quad(self):
        N, G = self.nCell, self.nGCell
        pts, wj = np.zeros((N, G, 3)), np.zeros((N, G))
        _chk(lib().orc_get_cell_quad(self.h, P(pts), P(wj)))
        return pts, wj

    def vr_coeffs(self):
        N, M, K = self.nCell, self.M, self.maxFC
        AinvB = np.zeros((N, K + 1, M, M))
        Ainvb = np.zeros((N, K, M))
        _chk(lib().orc_get_vr_coeffs(self.h, P(AinvB), P(Ainvb)))
        return AinvB, Ainvb

    # ---- fields
    def cell_average(self, fieldId, params, nVars):
        u = np.zeros((self.nCell, nVars))
        prm = np.ascontiguousarray(params, np.float64)
        _chk(lib().orc_cell_average(self.h, C.c_int(fieldId), P(prm), C.c_int(nVars), P(u)))
        return u

    def rec_error_l1(self, fieldId, params, u, uRec):
        prm = np.ascontiguousarray(params, np.float64)
        return lib().orc_rec_error_l1(self.h, C.c_int(fieldId), P(prm), P(np.ascontiguousarray(u)),
                                      P(np.ascontiguousarray(uRec)))

    # ---- VR
    def vr_iter(self, uRec, u):
        nVars = u.shape[1]
        out = np.zeros_like(uRec)
        _chk(lib().orc_vr_iter(self.h, C.c_int(nVars), P(np.ascontiguousarray(uRec)), P(np.ascontiguousarray(u)),
                               P(out)))
        return out

    def vr_2nd(self, u):
        """Second-order substitute (settings subs2ndOrder): uRec from the means."""
        nVars = u.shape[1]
        out = np.zeros((self.nCell, nVars, self.M))
        _chk(lib().orc_vr_2nd(self.h, C.c_int(nVars), P(np.ascontiguousarray(u)), P(out)))
        return out

    def vr_sor(self, uRec, u, scanOrder=None):
        """One in-place SOR sweep (cells in scanOrder, None = natural order); returns the new uRec."""
        nVars = u.shape[1]
        out = np.array(uRec, dtype=np.float64, copy=True)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        _chk(lib().orc_vr_sor(self.h, C.c_int(nVars), P(out), P(np.ascontiguousarray(u)), P(so)))
        return out

    def vr_iter_diff(self, uRecDiff, uRec=None, u=None):
        nVars = uRecDiff.shape[1]
        out = np.zeros_like(uRecDiff)
        _chk(lib().orc_vr_iter_diff(self.h, C.c_int(nVars), None if uRec is None else P(np.ascontiguousarray(uRec)),
                                    None if u is None else P(np.ascontiguousarray(u)),
                                    P(np.ascontiguousarray(uRecDiff)), P(out)))
        return out

    # ---- boundary conditions
    def set_bc(self, types, values=None, gamma=1.4, Rgas=1.0):
        """Per-zone EulerBCType codes and 5-vectors (far-field / inflow states; the wall
        temperature of BCWallIsothermal; p0, T0 and the flow direction of BCInPsTs)."""
        t = np.ascontiguousarray(types, np.int32)
        v = None if values is None else np.ascontiguousarray(values, np.float64).reshape(len(t), 5)
        self._bc = (t, v)
        _chk(lib().orc_set_bc(self.h, C.c_int(len(t)), P(t), P(v), C.c_double(gamma), C.c_double(Rgas)))

    def set_vr_dirichlet(self, values):
        """Boundary point values [nFace, nGFace, nVars] for the VR sweep (boundary faces read)."""
        v = np.ascontiguousarray(values, np.float64)
        _chk(lib().orc_set_vr_dirichlet(self.h, C.c_int(v.shape[2]), P(v)))

    # ---- Euler
    @staticmethod
    def euler_settings(gamma=1.4, rsType=1, rsMeanValueEig=0, rsFixScale=1.0, rsIncFScale=1.0, viscous=0,
                       muModel=1, Rgas=1.0, muGas=1.0, prGas=0.72, TRef=273.15, CSutherland=110.4, useRoeJacobian=0):
        return EulerSettings(gamma, rsType, rsMeanValueEig, rsFixScale, rsIncFScale, viscous, muModel, Rgas, muGas,
                             prGas, TRef, CSutherland, int(useRoeJacobian))

    def eval_dt(self, es, u, CFL, maxDt=1e100, useLocalDt=True):
        N, F = self.nCell, self.nFace
        dt, lam, dlc = np.zeros(N), np.zeros((F, 4)), np.zeros(N)
        dmin = C.c_double()
        _chk(lib().orc_eval_dt(self.h, C.byref(es), P(np.ascontiguousarray(u)), C.c_double(CFL), C.c_double(maxDt),
                               C.c_int(int(useLocalDt)), P(dt), P(lam), P(dlc), C.byref(dmin)))
        return dt, lam, dlc, dmin.value

    def eval_rhs(self, es, u, uRec=None, deltaLambdaCell=None, roeLambdas=False):
        """EvaluateRHS; with roeLambdas also the face-averaged Riemann eigenvalues
        lambdaFace0/123/4 [nFace, 3] that useRoeJacobian keeps (EvaluateRHS.hxx:527-544)."""
        rhs = np.zeros((self.nCell, 5))
        lr = np.zeros((self.nFace, 3)) if roeLambdas else None
        _chk(lib().orc_eval_rhs_lam(self.h, C.byref(es), P(np.ascontiguousarray(u)),
                                    None if uRec is None else P(np.ascontiguousarray(uRec)),
                                    None if deltaLambdaCell is None else P(np.ascontiguousarray(deltaLambdaCell)),
                                    P(rhs), P(lr)))
        return (rhs, lr) if roeLambdas else rhs

    def zone_force(self, es, u, uRec, deltaLambdaCell, zone):
        """Force on the boundary faces of a zone (the wall force of a drag functional), [3]."""
        f = np.zeros(3)
        _chk(lib().orc_zone_force(self.h, C.byref(es), P(np.ascontiguousarray(u)),
                                  None if uRec is None else P(np.ascontiguousarray(uRec)),
                                  None if deltaLambdaCell is None else P(np.ascontiguousarray(deltaLambdaCell)),
                                  C.c_int(zone), P(f)))
        return f

    def update_sgs_transpose(self, es, alphaDiag, rhs, u, uInc, uIncNew, jdiag, lam, forward, gsUpdate, uIncIsZero,
                             scanOrder=None):
        """A sweep of the transposed scalar LU-SGS matrix (the adjoint preconditioner)."""
        out = np.array(uIncNew, dtype=np.float64, copy=True)
        s = np.zeros(5)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        _chk(lib().orc_update_sgs_transpose(self.h, C.byref(es), C.c_double(alphaDiag), P(np.ascontiguousarray(rhs)),
                                            P(np.ascontiguousarray(u)), P(np.ascontiguousarray(uInc)), P(out),
                                            P(np.ascontiguousarray(jdiag)), P(np.ascontiguousarray(lam)),
                                            C.c_int(int(forward)), C.c_int(int(gsUpdate)), C.c_int(int(uIncIsZero)),
                                            P(so), P(s)))
        return out, s

    def update_sgs_block_transpose(self, es, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe, forward,
                                   gsUpdate, uIncIsZero, scanOrder=None):
        """A sweep of the transposed block LU-SGS matrix (the block adjoint preconditioner)."""
        out = np.array(uIncNew, dtype=np.float64, copy=True)
        s = np.zeros(5)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        _chk(lib().orc_update_sgs_block_transpose(
            self.h, C.byref(es), C.c_double(alphaDiag), P(np.ascontiguousarray(rhs)), P(np.ascontiguousarray(u)),
            P(np.ascontiguousarray(uInc)), P(out), P(np.ascontiguousarray(jinv)), P(np.ascontiguousarray(lam)),
            None if lamRoe is None else P(np.ascontiguousarray(lamRoe)), C.c_int(int(forward)),
            C.c_int(int(gsUpdate)), C.c_int(int(uIncIsZero)), P(so), P(s)))
        return out, s

    def lusgs_init_block(self, es, dTau, dt, alphaDiag, u, lam, lamRoe=None):
        """LUSGSMatrixInit with useRoeJacobian: (JDiag [N, 5, 5], its inverse [N, 5, 5])."""
        jd, ji = np.zeros((self.nCell, 5, 5)), np.zeros((self.nCell, 5, 5))
        _chk(lib().orc_lusgs_init_block(self.h, C.byref(es), P(np.ascontiguousarray(dTau)), C.c_double(dt),
                                        C.c_double(alphaDiag), P(np.ascontiguousarray(u)), P(np.ascontiguousarray(lam)),
                                        None if lamRoe is None else P(np.ascontiguousarray(lamRoe)), P(jd), P(ji)))
        return jd, ji

    def update_sgs_block(self, es, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe, forward, gsUpdate,
                         uIncIsZero, scanOrder=None):
        out = np.array(uIncNew, dtype=np.float64, copy=True)
        s = np.zeros(5)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        _chk(lib().orc_update_sgs_block(self.h, C.byref(es), C.c_double(alphaDiag), P(np.ascontiguousarray(rhs)),
                                        P(np.ascontiguousarray(u)), P(np.ascontiguousarray(uInc)), P(out),
                                        P(np.ascontiguousarray(jinv)), P(np.ascontiguousarray(lam)),
                                        None if lamRoe is None else P(np.ascontiguousarray(lamRoe)),
                                        C.c_int(int(forward)), C.c_int(int(gsUpdate)), C.c_int(int(uIncIsZero)), P(so),
                                        P(s)))
        return out, s

    def lusgs_init(self, dTau, dt, alphaDiag, lam):
        jd = np.zeros(self.nCell)
        _chk(lib().orc_lusgs_init(self.h, P(np.ascontiguousarray(dTau)), C.c_double(dt), C.c_double(alphaDiag),
                                  P(np.ascontiguousarray(lam)), P(jd)))
        return jd

    def update_sgs(self, es, alphaDiag, rhs, u, uInc, uIncNew, jdiag, lam, forward, gsUpdate, uIncIsZero,
                   scanOrder=None):
        out = np.array(uIncNew, dtype=np.float64, copy=True)
        s = np.zeros(5)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        _chk(lib().orc_update_sgs(self.h, C.byref(es), C.c_double(alphaDiag), P(np.ascontiguousarray(rhs)),
                                  P(np.ascontiguousarray(u)), P(np.ascontiguousarray(uInc)), P(out),
                                  P(np.ascontiguousarray(jdiag)), P(np.ascontiguousarray(lam)), C.c_int(int(forward)),
                                  C.c_int(int(gsUpdate)), C.c_int(int(uIncIsZero)), P(so), P(s)))
        return out, s

    def ssprk3(self, es, u, uRec, nStep, CFL, nRec=1, dtFixed=0.0):
        u = np.array(u, dtype=np.float64, copy=True)
        uRec = np.array(uRec, dtype=np.float64, copy=True)
        _chk(lib().orc_ssprk3(self.h, C.byref(es), P(u), P(uRec), C.c_int(nStep), C.c_double(CFL), C.c_int(nRec),
                              C.c_double(dtFixed)))
        return u, uRec

    def implicit_euler_step(self, es, u, uRec, dt, CFL=10.0, nRec=1, sgsIter=0, gsUpdate=1, maxIter=20, minIter=5,
                            tol=1e-10, incrementRelax=0.9, refRho=1.0, refP=1.0, scanOrder=None, stats=False):
        """Implicit-Euler dual-time step; returns (u, uRec, nIter, res, resRel, history), plus
        (number of positivity-limited cell increments, smallest alpha) when stats."""
        u = np.array(u, dtype=np.float64, copy=True)
        uRec = np.array(uRec, dtype=np.float64, copy=True)
        st = ImplicitSettings(dt, CFL, nRec, sgsIter, gsUpdate, maxIter, minIter, tol, 0, incrementRelax, refRho, refP)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        n = C.c_int()
        res, rel, hist = np.zeros(5), np.zeros(5), np.zeros(max(maxIter, 1))
        pps = np.zeros(2)
        _chk(lib().orc_implicit_euler_step(self.h, C.byref(es), C.byref(st), P(u), P(uRec), P(so), C.byref(n), P(res),
                                           P(rel), P(hist), P(pps)))
        if stats:
            return u, uRec, n.value, res, rel, hist[:n.value], (int(pps[0]), pps[1])
        return u, uRec, n.value, res, rel, hist[:n.value]

    def implicit_bdf_step(self, es, u, uRec, dt, xPrevs=(), CFL=10.0, nRec=1, sgsIter=0, gsUpdate=1, maxIter=20,
                          minIter=5, tol=1e-10, incrementRelax=0.9, refRho=1.0, refP=1.0, scanOrder=None):
        """ImplicitBDFDualTimeStep::Step with the earlier states xPrevs (newest first, up to 3):
        the (len(xPrevs) + 1)-step BDF formula.  Returns (u, uRec, nIter, res, resRel, history)."""
        u = np.array(u, dtype=np.float64, copy=True)
        uRec = np.array(uRec, dtype=np.float64, copy=True)
        st = ImplicitSettings(dt, CFL, nRec, sgsIter, gsUpdate, maxIter, minIter, tol, 0, incrementRelax, refRho, refP)
        so = None if scanOrder is None else np.ascontiguousarray(scanOrder, np.int64)
        xp = np.ascontiguousarray(np.stack(xPrevs), np.float64) if len(xPrevs) else None
        n = C.c_int()
        res, rel, hist = np.zeros(5), np.zeros(5), np.zeros(max(maxIter, 1))
        _chk(lib().orc_implicit_bdf_step(self.h, C.byref(es), C.byref(st), C.c_int(len(xPrevs)), P(xp), P(u),
                                         P(uRec), P(so), C.byref(n), P(res), P(rel), P(hist), None))
        return u, uRec, n.value, res, rel, hist[:n.value]

    def model_rhs(self, u, uRec=None, ax=1.0, ay=0.0, sigma=0.0):
        """ModelEvaluator::EvaluateRHS (2-D scalar advection-diffusion of each column of u)."""
        nVars = u.shape[1]
        out = np.zeros((self.nCell, nVars))
        _chk(lib().orc_model_rhs(self.h, C.c_int(nVars), C.c_double(ax), C.c_double(ay), C.c_double(sigma),
                                 P(np.ascontiguousarray(u, np.float64)),
                                 None if uRec is None else P(np.ascontiguousarray(uRec, np.float64)), P(out)))
        return out

    def lusgs_matvec(self, es, alphaDiag, u, x, jdiag, lam):
        """LUSGSMatrixVec: the scalar-Jacobian LU-SGS matrix applied to x."""
        out = np.zeros((self.nCell, 5))
        _chk(lib().orc_lusgs_matvec(self.h, C.byref(es), C.c_double(alphaDiag), P(np.ascontiguousarray(u)),
                                    P(np.ascontiguousarray(x)), P(np.ascontiguousarray(jdiag)),
                                    P(np.ascontiguousarray(lam)), P(out)))
        return out
