"""EulerEvaluator on the device, with the reference's method names.

Mirrors ``DNDS::Euler::EulerEvaluator<model>`` for the NS / NS_3D models
(src/Euler/EulerEvaluator.hpp:76): EvaluateDt (:359), EvaluateRHS (:399),
LUSGSMatrixInit (:427), UpdateLUSGSForward/Backward (:481-502), UpdateSGS (:513),
EvaluateNorm (:613).  Like the reference, the evaluator keeps lambdaFace* and
deltaLambdaCell from the last EvaluateDt as members.

The laminar viscous terms are on with ``settings["viscous"] = True``; the gas then
follows ``settings["idealGasProperty"]`` (gamma, Rgas, muGas, prGas, TRef, CSutherland,
muModel: 0 constant, 1 Sutherland, 2 constant nu; defaults as
src/Euler/EulerEvaluatorSettings.hpp:291-300).  Without it the residual is the
inviscid one (the reference's RHS_Ignore_Viscosity flag, EvaluateRHS.hxx:78).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .cfv import RIEMANN_TYPES, VariationalReconstruction, _ptr


def _check_jacobi_code(code):
    """linearSolverControl.jacobiCode (EulerSolver.hpp:703): 0 Jacobi, 1 Gauss-Seidel; 2 (ILU)
    is not provided by this path."""
    if code not in (0, 1):
        raise ValueError(f"jacobiCode {code} is not supported: 0 (Jacobi) or 1 (Gauss-Seidel); "
                         "2 (ILU) is not provided")


class JacobianDiagBlock:
    """5x5 diagonal blocks of the LU-SGS matrix and their inverses, [25, nCell] each (row-major
    block per cell), the block mode of the reference's JacobianDiagBlock
    (src/Euler/EulerJacobian.hpp:42-260) that useRoeJacobian needs.  ``value`` is the block,
    ``invert`` its inverse (GetInvert, :153-190); dtype float64, or float32 for the FP32 sweep."""

    def __init__(self, nCell, device, dtype=torch.float64):
        self.value = torch.zeros((25, nCell), dtype=dtype, device=device)
        self.invert = torch.zeros((25, nCell), dtype=dtype, device=device)

    def to(self, dtype):
        out = JacobianDiagBlock.__new__(JacobianDiagBlock)
        out.value, out.invert = self.value.to(dtype), self.invert.to(dtype)
        return out

    def getBlock(self, iCell):
        return self.value[:, iCell].reshape(5, 5)


class ImplicitBDFDualTimeStep:
    """The reference's k-step BDF stepper (src/Solver/ODE.hpp:362-596) over an EulerEvaluator:
    it keeps the k - 1 earlier states on the device and starts at first order, raising the
    order by one per step up to k (kCurrent = number of stored states + 1)."""

    def __init__(self, ev, k: int = 2):
        if not 1 <= k <= 4:
            raise ValueError("k must be 1..4")
        self.ev, self.k = ev, k
        self.xPrevs = []
        self.dtPrev = None

    def Step(self, u, uRec, dt, **settings):
        if self.xPrevs and abs(self.dtPrev - dt) > 1e-8 * dt:
            raise ValueError("the BDF stepper takes a constant dt (ODE.hpp:432)")
        xLast = u.clone()
        out = self.ev.ImplicitBDFStep(u, uRec, dt, self.xPrevs, **settings)
        if self.k > 1:
            self.xPrevs = ([xLast] + self.xPrevs)[: self.k - 1]
        self.dtPrev = dt
        return out


class EulerEvaluator:
    nVars = 5

    def __init__(self, vfv: VariationalReconstruction, settings: dict | None = None):
        s = dict(gamma=1.4, riemannSolverType="Roe", rsMeanValueEig=0, rsFixScale=1.0, rsIncFScale=1.0,
                 viscous=False, useRoeJacobian=False)
        gas = dict(Rgas=1.0, muGas=1.0, prGas=0.72, TRef=273.15, CSutherland=110.4, muModel=1)
        if settings:
            s.update(settings)
            if "idealGasProperty" in settings:
                igp = settings["idealGasProperty"]
                s["gamma"] = igp.get("gamma", s["gamma"])
                gas.update({k: igp[k] for k in gas if k in igp})
        s["idealGasProperty"] = dict(gas, gamma=s["gamma"])
        rs = s["riemannSolverType"]
        self.settings = s
        self.es = _lib.EulerSettings(float(s["gamma"]), RIEMANN_TYPES[rs] if isinstance(rs, str) else int(rs),
                                     int(s["rsMeanValueEig"]), float(s["rsFixScale"]), float(s["rsIncFScale"]),
                                     int(bool(s["viscous"])), int(gas["muModel"]), float(gas["Rgas"]),
                                     float(gas["muGas"]), float(gas["prGas"]), float(gas["TRef"]),
                                     float(gas["CSutherland"]), int(bool(s["useRoeJacobian"])))
        bcTypes = s.get("bcTypes")
        self._bc = None
        if bcTypes is not None:
            # boundary weights of the VR functional follow the boundary types, as the
            # reference's constructor sets them (EulerEvaluator.hpp:138-152)
            from . import bc as _bc
            codes = [_bc.bc_code(t) for t in bcTypes]
            w = _bc.euler_dir_weights(codes, vfv.settings["maxOrder"])
            vfv.ConstructBaseAndWeight(lambda z, p: w[z, p])
        vfv.construct()
        self.vfv = vfv
        self.ctx = vfv.ctx
        if bcTypes is not None:
            t = np.ascontiguousarray(codes, np.int32)
            vals = s.get("bcValues")
            v = None if vals is None else np.ascontiguousarray(vals, np.float64).reshape(len(t), 5)
            self._bc = (t, v)
            check(lib().dndsr_set_bc_gas(self.ctx.h, len(t), t.ctypes.data_as(_lib.i32p),
                                         None if v is None else v.ctypes.data_as(_lib.dp), float(s["gamma"]),
                                         float(gas["Rgas"])))
        dev = self.ctx.torch_device
        self.lambdaFace = torch.zeros((4, vfv.nFace), dtype=torch.float64, device=dev)
        # lambdaFace0/123/4 of the Roe Jacobian, written by EvaluateRHS when useRoeJacobian
        self.lambdaFaceRoe = torch.zeros((3, vfv.nFace), dtype=torch.float64, device=dev)
        self.deltaLambdaCell = torch.zeros(vfv.nCell, dtype=torch.float64, device=dev)
        self.dtMinDev = torch.zeros(1, dtype=torch.float64, device=dev)

    # EvaluateDt(dt, u, uRec, CFL, dtMinall&, MaxDt, UseLocaldt, t)
    def EvaluateDt(self, dt, u, uRec=None, CFL=1.0, MaxDt=1e100, UseLocaldt=True, t=0.0, sync_min=True):
        dmin = C.c_double()
        check(lib().dndsr_eval_dt(self.ctx.h, C.byref(self.es), _ptr(u), float(CFL), float(MaxDt), int(UseLocaldt),
                                  _ptr(dt), _ptr(self.lambdaFace), _ptr(self.deltaLambdaCell), _ptr(self.dtMinDev),
                                  C.byref(dmin) if sync_min else None))
        return dmin.value if sync_min else None

    # EvaluateRHS(rhs, JSource, u, uRecUnlim, uRec, uRecBeta, cellRHSAlpha, onlyOnHalfAlpha, t, flags)
    def EvaluateRHS(self, rhs, JSource, u, uRecUnlim, uRec=None, uRecBeta=None, cellRHSAlpha=None,
                    onlyOnHalfAlpha=False, t=0.0, flags=0, check_finite=True):
        """check_finite synchronises and raises when a cell's residual is not finite (the
        reference asserts, EvaluateRHS.hxx:387-398); without it the check waits for the next
        Context.sync()."""
        rec = uRec if uRec is not None else uRecUnlim
        check(lib().dndsr_eval_rhs(self.ctx.h, C.byref(self.es), _ptr(u), _ptr(rec), _ptr(self.deltaLambdaCell),
                                   _ptr(rhs)))
        if self.settings["useRoeJacobian"]:  # EvaluateRHS.hxx:527-544
            check(lib().dndsr_eval_rhs_roe_lambda(self.ctx.h, C.byref(self.es), _ptr(u), _ptr(rec),
                                                  _ptr(self.deltaLambdaCell), _ptr(self.lambdaFaceRoe)))
        if check_finite:
            self.ctx.sync()

    # LUSGSMatrixInit(JDiag, JSource, dTau, dt, alphaDiag, u, uRec, jacobianCode, t)
    def BuildJacobianDiagBlock(self):
        """The JDiag container of useRoeJacobian (block mode)."""
        return JacobianDiagBlock(self.vfv.nCell, self.lambdaFace.device)

    def LUSGSMatrixInit(self, JDiag, JSource, dTau, dt, alphaDiag, u=None, uRec=None, jacobianCode=0, t=0.0):
        """Scalar JDiag (a [nCell] tensor), or with useRoeJacobian a JacobianDiagBlock filled with the
        Roe-Jacobian blocks and their inverses (EulerEvaluator.hxx:37-159); the block mode uses the
        face eigenvalues of the last EvaluateRHS."""
        if jacobianCode != 0:
            raise ValueError("jacobianCode must be 0")
        if self.settings["useRoeJacobian"]:
            if not isinstance(JDiag, JacobianDiagBlock) or u is None:
                raise ValueError("useRoeJacobian needs a JacobianDiagBlock (BuildJacobianDiagBlock) and u")
            check(lib().dndsr_lusgs_init_block(self.ctx.h, C.byref(self.es), _ptr(dTau), float(dt),
                                               float(alphaDiag), _ptr(u), _ptr(self.lambdaFace),
                                               _ptr(self.lambdaFaceRoe), _ptr(JDiag.value), _ptr(JDiag.invert)))
            return
        check(lib().dndsr_lusgs_init(self.ctx.h, _ptr(dTau), float(dt), float(alphaDiag), _ptr(self.lambdaFace),
                                     _ptr(JDiag)))

    # UpdateSGS(alphaDiag, t, rhs, u, uInc, uIncNew, JDiag, forward, gsUpdate, sumInc&, uIncIsZero)
    def UpdateSGS(self, alphaDiag, t, rhs, u, uInc, uIncNew, JDiag, forward, gsUpdate, uIncIsZero=False,
                  lambdaFace=None):
        s = np.zeros(5)
        lam = self.lambdaFace if lambdaFace is None else lambdaFace
        if isinstance(JDiag, JacobianDiagBlock):
            if rhs.dtype == torch.float32:
                lr = self.lambdaFaceRoe.float()
                check(lib().dndsr_update_sgs_block_f32(
                    self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u), _ptr(uInc), _ptr(uIncNew),
                    _ptr(JDiag.invert), _ptr(lam), _ptr(lr), int(forward), int(gsUpdate), int(uIncIsZero),
                    s.ctypes.data_as(_lib.dp)))
            else:
                check(lib().dndsr_update_sgs_block(
                    self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u), _ptr(uInc), _ptr(uIncNew),
                    _ptr(JDiag.invert), _ptr(lam), _ptr(self.lambdaFaceRoe), int(forward), int(gsUpdate),
                    int(uIncIsZero), s.ctypes.data_as(_lib.dp)))
            return s
        if rhs.dtype == torch.float32:
            check(lib().dndsr_update_sgs_f32(self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u),
                                             _ptr(uInc), _ptr(uIncNew), _ptr(JDiag), _ptr(lam), int(forward),
                                             int(gsUpdate), int(uIncIsZero), s.ctypes.data_as(_lib.dp)))
        else:
            check(lib().dndsr_update_sgs(self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u),
                                         _ptr(uInc), _ptr(uIncNew), _ptr(JDiag), _ptr(lam), int(forward),
                                         int(gsUpdate), int(uIncIsZero), s.ctypes.data_as(_lib.dp)))
        return s

    def UpdateSGSTranspose(self, alphaDiag, rhs, u, uInc, uIncNew, JDiag, forward, gsUpdate, uIncIsZero=False):
        """A sweep of the transpose of the LU-SGS matrix that UpdateSGS sweeps with the same JDiag
        (scalar, or a JacobianDiagBlock with the face eigenvalues of the last EvaluateRHS), in the
        same colour order, FP64: the preconditioner of the steady adjoint."""
        s = np.zeros(5)
        if isinstance(JDiag, JacobianDiagBlock):
            check(lib().dndsr_update_sgs_block_transpose(
                self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u), _ptr(uInc), _ptr(uIncNew),
                _ptr(JDiag.invert), _ptr(self.lambdaFace), _ptr(self.lambdaFaceRoe), int(forward), int(gsUpdate),
                int(uIncIsZero), s.ctypes.data_as(_lib.dp)))
            return s
        check(lib().dndsr_update_sgs_transpose(self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u),
                                               _ptr(uInc), _ptr(uIncNew), _ptr(JDiag), _ptr(self.lambdaFace),
                                               int(forward), int(gsUpdate), int(uIncIsZero),
                                               s.ctypes.data_as(_lib.dp)))
        return s

    def SGSTransposeSymmetric(self, alphaDiag, rhs, u, JDiag, nSweep, gsUpdate, out=None, work=None):
        """nSweep symmetric sweeps of the transposed LU-SGS matrix from zero increments in one call
        (dndsr_sgs_transpose_symmetric): the result of nSweep forward/backward pairs of
        UpdateSGSTranspose, without a host synchronisation per sweep.  Returns out."""
        out = torch.empty_like(rhs) if out is None else out
        work = torch.empty_like(rhs) if work is None else work
        blk = isinstance(JDiag, JacobianDiagBlock)
        check(lib().dndsr_sgs_transpose_symmetric(
            self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(rhs), _ptr(u), None if blk else _ptr(JDiag),
            _ptr(JDiag.invert) if blk else None, _ptr(self.lambdaFace), _ptr(self.lambdaFaceRoe) if blk else None,
            int(nSweep), int(gsUpdate), _ptr(work), _ptr(out)))
        return out

    def SetBCValues(self, bcValues):
        """Replaces the per-zone boundary states (5 conservative values per zone) of the
        evaluator's boundary set; the zone types stay."""
        if self._bc is None:
            raise ValueError("SetBCValues: the evaluator was built without bcTypes")
        t = self._bc[0]
        v = np.ascontiguousarray(bcValues, np.float64).reshape(len(t), 5)
        self._bc = (t, v)
        check(lib().dndsr_set_bc_gas(self.ctx.h, len(t), t.ctypes.data_as(_lib.i32p), v.ctypes.data_as(_lib.dp),
                                     float(self.settings["gamma"]), float(self.settings["idealGasProperty"]["Rgas"])))

    # deprecated LU-SGS pair (EulerEvaluator.hxx:376-529): Forward writes D^-1 (r - L x_new),
    # Backward adds D^-1 (-U x); with separate uInc they reduce to block Jacobi.
    def UpdateLUSGSForward(self, alphaDiag, t, rhs, u, uInc, JDiag, uIncNew):
        return self.UpdateSGS(alphaDiag, t, rhs, u, uInc, uIncNew, JDiag, True, True, True)

    def UpdateLUSGSBackward(self, alphaDiag, t, rhs, u, uInc, JDiag, uIncNew):
        tmp = uIncNew.clone()
        return self.UpdateSGS(alphaDiag, t, rhs, u, tmp, uIncNew, JDiag, False, True, False)

    # FixUMaxFilter(u), EulerEvaluator.hxx:1318-1323: the reference returns at once
    def FixUMaxFilter(self, u):
        return None

    # EvaluateNorm(res, rhs, P, volWise, average), EulerEvaluator.hxx:1394-1440
    def EvaluateNorm(self, rhs, P=1, volWise=False, average=False):
        vol = torch.as_tensor(self.vfv.cell_metrics()["vol"], device=rhs.device)
        a = rhs.abs().double()
        if P < 3:
            w = vol if volWise else torch.ones_like(vol)
            res = (a.pow(P) * w).sum(dim=1).pow(1.0 / P)
            if average:
                res = res / (w.sum() + 1e-200)
            return res.cpu().numpy()
        return a.max(dim=1).values.cpu().numpy()

    # LUSGSMatrixVec(alphaDiag, t, u, uInc, JDiag, AuInc), EulerEvaluator.hpp:449-455
    def LUSGSMatrixVec(self, alphaDiag, t, u, uInc, JDiag, AuInc, lambdaFace=None):
        if isinstance(JDiag, JacobianDiagBlock):
            raise ValueError("LUSGSMatrixVec is provided for the scalar Jacobian only")
        lam = self.lambdaFace if lambdaFace is None else lambdaFace
        check(lib().dndsr_lusgs_matvec(self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(u), _ptr(uInc),
                                       _ptr(JDiag), _ptr(lam), _ptr(AuInc)))

    def GMRES(self, alphaDiag, u, JDiag, b, x, nGmresSpace=10, nGmresIter=2, sgsIter=0, jacobiCode=1,
              fp32Precond=False, lambdaFace=None):
        """GMRES_LeftPreconditioned::solve (Linear.hpp:47-131) on the LU-SGS matrix with the
        LU-SGS preconditioner (FP32 sweeps when fp32Precond); x: initial guess -> solution.
        Returns (last preconditioned residual norm, |M b|)."""
        if isinstance(JDiag, JacobianDiagBlock):
            raise ValueError("GMRES is provided for the scalar Jacobian only")
        lam = self.lambdaFace if lambdaFace is None else lambdaFace
        _check_jacobi_code(jacobiCode)
        r = np.zeros(2)
        check(lib().dndsr_gmres_lusgs(self.ctx.h, C.byref(self.es), float(alphaDiag), _ptr(u), _ptr(JDiag),
                                      _ptr(lam), _ptr(b), _ptr(x), int(nGmresSpace), int(nGmresIter), int(sgsIter),
                                      int(jacobiCode), int(bool(fp32Precond)), r.ctypes.data_as(_lib.dp)))
        return r[0], r[1]

    def ImplicitEulerStep(self, u, uRec, dt, CFL=10.0, nRec=1, sgsIter=0, jacobiCode=1, maxIter=20, minIter=5,
                          tol=1e-10, fp32Sweeps=False, incrementPPRelax=0.9, refRho=1.0, refP=1.0, gmresCode=0,
                          nGmresSpace=10, nGmresIter=2):
        """One implicit-Euler dual-time step on the device (ImplicitEulerDualTimeStep::Step,
        ODE.hpp:80-98, with EulerSolver's fdtau / frhs / fsolve / fincrement / fstop,
        EulerSolver.hxx:645-830, 1206).  Defaults follow EulerSolver's config defaults
        (EulerSolver.hpp:171-760).  fp32Sweeps runs the LU-SGS sweeps (the solver, or the GMRES
        preconditioner when gmresCode is 1 or 2) in FP32 while the residual, reconstruction,
        Krylov vectors and update stay FP64.  The increment is positivity-limited and compressed
        as the reference's fincrement (PPStats() gives the counts).  Returns (nIter, res, resRel,
        history)."""
        _check_jacobi_code(jacobiCode)
        s = _lib.ImplicitSettings(float(dt), float(CFL), int(nRec), int(sgsIter), int(jacobiCode), int(maxIter),
                                  int(minIter), float(tol), int(bool(fp32Sweeps)), float(incrementPPRelax),
                                  float(refRho), float(refP), int(gmresCode), int(nGmresSpace), int(nGmresIter))
        n = C.c_int()
        res, rel, hist = np.zeros(5), np.zeros(5), np.zeros(max(int(maxIter), 1))
        check(lib().dndsr_implicit_euler_step(self.ctx.h, C.byref(self.es), C.byref(s), _ptr(u), _ptr(uRec),
                                              C.byref(n), res.ctypes.data_as(_lib.dp), rel.ctypes.data_as(_lib.dp),
                                              hist.ctypes.data_as(_lib.dp)))
        return n.value, res, rel, hist[:n.value]

    def ImplicitBDFStep(self, u, uRec, dt, xPrevs=(), CFL=10.0, nRec=1, sgsIter=0, jacobiCode=1, maxIter=20,
                        minIter=5, tol=1e-10, fp32Sweeps=False, incrementPPRelax=0.9, refRho=1.0, refP=1.0,
                        gmresCode=0, nGmresSpace=10, nGmresIter=2):
        """One BDF dual-time step (ImplicitBDFDualTimeStep::Step, src/Solver/ODE.hpp:427-477) with
        the earlier states xPrevs (device [5, nCell] tensors, newest first, up to three): the
        (len(xPrevs) + 1)-step formula.  Settings and return value as ImplicitEulerStep."""
        _check_jacobi_code(jacobiCode)
        if len(xPrevs) > 3:
            raise ValueError("at most three earlier states (BDF4)")
        s = _lib.ImplicitSettings(float(dt), float(CFL), int(nRec), int(sgsIter), int(jacobiCode), int(maxIter),
                                  int(minIter), float(tol), int(bool(fp32Sweeps)), float(incrementPPRelax),
                                  float(refRho), float(refP), int(gmresCode), int(nGmresSpace), int(nGmresIter))
        ptrs = (C.c_void_p * max(len(xPrevs), 1))(*[x.data_ptr() for x in xPrevs])
        n = C.c_int()
        res, rel, hist = np.zeros(5), np.zeros(5), np.zeros(max(int(maxIter), 1))
        check(lib().dndsr_implicit_bdf_step(self.ctx.h, C.byref(self.es), C.byref(s), len(xPrevs),
                                            C.cast(ptrs, C.POINTER(C.c_void_p)), _ptr(u), _ptr(uRec), C.byref(n),
                                            res.ctypes.data_as(_lib.dp), rel.ctypes.data_as(_lib.dp),
                                            hist.ctypes.data_as(_lib.dp)))
        return n.value, res, rel, hist[:n.value]

    def PPStats(self):
        """(cell increments limited by the positivity limiter, smallest alpha) of the last
        ImplicitEulerStep (the solver's PPIncrementLimiter report, EulerSolver.hxx:708-718)."""
        n, a = C.c_int64(), C.c_double()
        check(lib().dndsr_implicit_pp_stats(self.ctx.h, C.byref(n), C.byref(a)))
        return n.value, a.value

    def SSPRK3(self, u, uRec, nStep, CFL, nRec=1, dtFixed=0.0):
        """Device-resident explicit SSP-RK3 loop (ODE.hpp:1437-1535)."""
        check(lib().dndsr_ssprk3(self.ctx.h, C.byref(self.es), _ptr(u), _ptr(uRec), int(nStep), float(CFL), int(nRec),
                                 float(dtFixed)))
