"""Discrete adjoint of the compact-FV residual as torch.autograd.Functions.

The reference has no adjoint; these wrap the device vector-Jacobian products
(dndsr_eval_rhs_vjp, dndsr_vr_iter_vjp) so that torch's autograd can chain them:

* ``EvaluateRHSFunction``: rhs = EvaluateRHS(u, uRec) with deltaLambdaCell fixed
  (EulerEvaluator::EvaluateRHS, src/Euler/EulerEvaluator.hpp:399-409);
* ``VRIterFunction``: one Jacobi VR sweep uRecNew = S(uRec, u)
  (VariationalReconstruction::DoReconstructionIter, VariationalReconstruction.hpp:984-992);
* ``ResidualFunction``: frhs = nRec Jacobi VR sweeps from uRec0 followed by
  EvaluateRHS (EulerSolver.hxx:147-235).  Only the inputs are kept: the backward
  pass recomputes the reconstruction chain, then runs the transposed operators.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .cfv import _ptr


def rhs_vjp(ev, u, uRec, deltaLambdaCell, rhsBar, need_rec=True):
    """(d rhs/d u)^T rhsBar, (d rhs/d uRec)^T rhsBar on the device."""
    uBar = torch.empty_like(u)
    recBar = torch.empty_like(uRec) if (uRec is not None and need_rec) else None
    check(lib().dndsr_eval_rhs_vjp(ev.ctx.h, C.byref(ev.es), _ptr(u), _ptr(uRec), _ptr(deltaLambdaCell),
                                   _ptr(rhsBar.contiguous()), _ptr(uBar), _ptr(recBar)))
    return uBar, recBar


def vr_vjp(vfv, recNewBar, mode=0, relax=-1.0, uRec=None, u=None):
    """Transpose of one VR sweep: returns (uRecBar, uBar) (uBar is None for mode 3 = IterDiff).
    uRec, u: the state the forward sweep read (needed on meshes with boundary faces)."""
    nVars = recNewBar.shape[0]
    recBar = torch.empty_like(recNewBar)
    uBar = None if mode == 3 else torch.empty((nVars, vfv.nCell), dtype=recNewBar.dtype, device=recNewBar.device)
    check(lib().dndsr_vr_iter_vjp(vfv.ctx.h, int(nVars), int(mode), float(relax),
                                  None if uRec is None else _ptr(uRec.contiguous()),
                                  None if u is None else _ptr(u.contiguous()), _ptr(recNewBar.contiguous()),
                                  _ptr(recBar), _ptr(uBar)))
    return recBar, uBar


def _eval_rhs(ev, u, uRec, dlc):
    rhs = torch.empty_like(u)
    check(lib().dndsr_eval_rhs(ev.ctx.h, C.byref(ev.es), _ptr(u), _ptr(uRec), _ptr(dlc), _ptr(rhs)))
    return rhs


def _vr_sweep(vfv, uRec, u, relax):
    out = torch.empty_like(uRec)
    check(lib().dndsr_vr_iter(vfv.ctx.h, int(u.shape[0]), _ptr(uRec), _ptr(out), _ptr(u), 0, float(relax)))
    return out


def _require_jacobi_frhs(vfv):
    """The residual's transpose chains transposed Jacobi sweeps (or, with subs2ndOrder, the
    transposed second-order substitute, which every sweep then is); with SORInstead the
    solver's frhs sweeps in place (multicolour SOR), which this chain does not differentiate."""
    if vfv.settings.get("SORInstead") and not vfv.settings.get("subs2ndOrder"):
        raise ValueError("the residual adjoint differentiates Jacobi VR sweeps; the settings select SORInstead "
                         "(the transpose of the in-place SOR sweep is not provided)")


class EvaluateRHSFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, uRec, ev, deltaLambdaCell):
        u, uRec = u.contiguous(), uRec.contiguous()
        ctx.ev = ev
        ctx.save_for_backward(u, uRec, deltaLambdaCell)
        return _eval_rhs(ev, u, uRec, deltaLambdaCell)

    @staticmethod
    def backward(ctx, rhsBar):
        u, uRec, dlc = ctx.saved_tensors
        uBar, recBar = rhs_vjp(ctx.ev, u, uRec, dlc, rhsBar, need_rec=ctx.needs_input_grad[1])
        return uBar, recBar, None, None


class VRIterFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uRec, u, vfv, relax=-1.0):
        ctx.vfv, ctx.relax = vfv, relax
        uRec, u = uRec.contiguous(), u.contiguous()
        ctx.save_for_backward(uRec, u)
        return _vr_sweep(vfv, uRec, u, relax)

    @staticmethod
    def backward(ctx, recNewBar):
        uRec, u = ctx.saved_tensors
        recBar, uBar = vr_vjp(ctx.vfv, recNewBar, 0, ctx.relax, uRec, u)
        return recBar, uBar, None, None


class ResidualFunction(torch.autograd.Function):
    """rhs = EvaluateRHS(u, S^nRec(uRec0, u)); backward recomputes the S chain."""

    @staticmethod
    def forward(ctx, u, uRec0, ev, nRec, deltaLambdaCell):
        _require_jacobi_frhs(ev.vfv)
        u, uRec0 = u.contiguous(), uRec0.contiguous()
        ctx.ev, ctx.nRec = ev, int(nRec)
        ctx.save_for_backward(u, uRec0, deltaLambdaCell)
        rec = uRec0
        for _ in range(ctx.nRec):
            rec = _vr_sweep(ev.vfv, rec, u, -1.0)
        return _eval_rhs(ev, u, rec, deltaLambdaCell)

    @staticmethod
    def backward(ctx, rhsBar):
        u, uRec0, dlc = ctx.saved_tensors
        ev = ctx.ev
        recs = [uRec0]
        for _ in range(ctx.nRec):
            recs.append(_vr_sweep(ev.vfv, recs[-1], u, -1.0))
        uBar, recBar = rhs_vjp(ev, u, recs[-1], dlc, rhsBar)
        for k in range(ctx.nRec - 1, -1, -1):
            recBar, ub = vr_vjp(ev.vfv, recBar, 0, -1.0, recs[k], u)
            uBar = uBar + ub
        return uBar, recBar, None, None, None


def residual(ev, u, uRec0, nRec=1, deltaLambdaCell=None):
    """Differentiable frhs(u, uRec0) on the device (deltaLambdaCell fixed; defaults to the
    evaluator's value from its last EvaluateDt)."""
    dlc = ev.deltaLambdaCell if deltaLambdaCell is None else deltaLambdaCell
    return ResidualFunction.apply(u, uRec0, ev, nRec, dlc)


def zone_force(ev, u, uRec, zone, deltaLambdaCell=None):
    """Force of the flow on boundary zone `zone` (momentum flux out of the domain through
    its faces, Riemann minus viscous flux) at (u, uRec); a length-3 numpy array."""
    dlc = ev.deltaLambdaCell if deltaLambdaCell is None else deltaLambdaCell
    out = np.zeros(3)
    check(lib().dndsr_zone_force(ev.ctx.h, C.byref(ev.es), _ptr(u.contiguous()), _ptr(uRec), _ptr(dlc), int(zone),
                                 out.ctypes.data_as(_lib.dp)))
    return out


def zone_force_vjp(ev, u, uRec, zone, forceBar, deltaLambdaCell=None, need_rec=True):
    """(d force / d u)^T forceBar and (d force / d uRec)^T forceBar on the device."""
    dlc = ev.deltaLambdaCell if deltaLambdaCell is None else deltaLambdaCell
    fb = np.ascontiguousarray(forceBar, np.float64)
    uBar = torch.empty_like(u)
    recBar = torch.empty_like(uRec) if (uRec is not None and need_rec) else None
    check(lib().dndsr_zone_force_vjp(ev.ctx.h, C.byref(ev.es), _ptr(u.contiguous()), _ptr(uRec), _ptr(dlc),
                                     int(zone), fb.ctypes.data_as(_lib.dp), _ptr(uBar), _ptr(recBar)))
    return uBar, recBar


# flux forms whose dissipation reads deltaLambdaCell: the entropy fixes of Roe_M6 / Roe_M8
# (Gas.hpp:878-1007, eig 6 and 8) and the HLLE+ dissipation factor (Gas.hpp:1150-1280)
_READS_DELTA_LAMBDA = (3, 16, 18, 21)


class SteadyAdjoint:
    """Discrete adjoint of a converged steady state on the device: the reference has none.

    At a steady state of the implicit solver (EulerSolver.hxx:645-830 marched with an
    infinite physical step) the state u and the reconstruction uRec satisfy

        R(u, uRec; p) = 0          (EvaluateRHS)
        uRec = S(uRec, u; p)       (a converged Jacobi VR sweep),

    where p are boundary values (with subs2ndOrder the reconstruction is the substitute
    uRec = G(u), which is S with dS/duRec = 0).  For a functional J(u, uRec) the multipliers
    lam (per cell, 5 values) and mu (per reconstruction coefficient) satisfy

        mu  = dJ/duRec + (dR/duRec)^T lam + (dS/duRec)^T mu
        0   = dJ/du    + (dR/du)^T lam    + (dS/du)^T mu,

    and dJ/dp = dJ/dp|explicit + lam . dR/dp + mu . dS/dp at fixed (u, uRec).  ``solve``
    solves them with the transposed operators (dndsr_eval_rhs_vjp, dndsr_vr_iter_vjp), either
    (method "fixed-point") by the transposed VR fixed point for mu and the transposed
    pseudo-time iteration lam += P^-T r, whose rate is the forward solver's, or (method
    "gmres", the default) by restarted right-preconditioned GMRES on the linear system in
    (lam, mu) with -P^-T on the lam block and -I on the mu block.  P^-T is (1 + sgsIter)
    symmetric sweeps of the transposed LU-SGS matrix built at u with the local pseudo time
    step of CFL: the scalar one (dndsr_update_sgs_transpose), or with block (the default
    when the evaluator has useRoeJacobian) the Roe-Jacobian block one
    (dndsr_update_sgs_block_transpose); the solution does not depend on P.  deltaLambdaCell is held at its value at u: flux forms whose dissipation reads it
    (Roe_M6, Roe_M8, HLLEP, HLLEP_V1) are refused unless rsFixScale is 0."""

    def __init__(self, ev, CFL=10.0, sgsIter=1, gsUpdate=True, block=None):
        if block is None:
            block = bool(ev.settings["useRoeJacobian"])
        if block and not ev.settings["useRoeJacobian"]:
            raise ValueError("the block preconditioner needs an evaluator with useRoeJacobian")
        if ev.es.rsType in _READS_DELTA_LAMBDA and ev.es.rsFixScale != 0.0:
            raise ValueError("the flux form reads deltaLambdaCell, which the steady adjoint holds fixed; "
                             "use rsFixScale = 0 or a flux form that does not read it")
        self.ev, self.CFL, self.sgsIter, self.gsUpdate = ev, float(CFL), int(sgsIter), bool(gsUpdate)
        self.block = bool(block)
        self.history = []

    def _prepare(self, u, uRec):
        ev = self.ev
        dTau = torch.empty(ev.vfv.nCell, dtype=torch.float64, device=u.device)
        ev.EvaluateDt(dTau, u, None, CFL=self.CFL)
        self.dlc = ev.deltaLambdaCell.clone()
        if self.block:  # the Roe-Jacobian blocks need the face eigenvalues of EvaluateRHS
            ev.EvaluateRHS(torch.empty_like(u), None, u, uRec)
            jd = ev.BuildJacobianDiagBlock()
            ev.LUSGSMatrixInit(jd, None, dTau, 1e100, 1.0, u, uRec)
            return jd
        jd = torch.empty_like(dTau)
        check(lib().dndsr_lusgs_init(ev.ctx.h, _ptr(dTau), 1e100, 1.0, _ptr(ev.lambdaFace), _ptr(jd)))
        return jd

    def _precond(self, r, u, jd):
        # (1 + sgsIter) symmetric transposed sweeps from zero, in one call
        return self.ev.SGSTransposeSymmetric(1.0, r.contiguous(), u, jd, self.sgsIter + 1, self.gsUpdate)

    def _apply(self, u, uRec, lam, mu):
        """The transposed steady system: ((dR/du)^T lam + (dS/du)^T mu,
        (dR/duRec)^T lam + (dS/duRec)^T mu - mu)."""
        uB, rB = rhs_vjp(self.ev, u, uRec, self.dlc, lam)
        rS, uS = vr_vjp(self.ev.vfv, mu, 0, -1.0, uRec, u)
        return uB + uS, rB + rS - mu

    def solve(self, u, uRec, uBarJ, recBarJ=None, tol=1e-11, maxIter=2000, lam0=None, mu0=None, method="gmres",
              restart=40):
        """Multipliers (lam, mu) for the functional gradients uBarJ = dJ/du ([5, nCell]) and
        recBarJ = dJ/duRec (uRec's shape, or None).  GMRES stops when the residual of the
        transposed system falls below tol times the norm of (uBarJ, recBarJ); the fixed-point
        iteration when the u-stationarity residual falls below tol times its first value and
        the last change of mu below tol times its norm.  history keeps the residual norm per
        iteration (for GMRES, its estimate).  Returns (lam, mu, nIter)."""
        u, uRec = u.contiguous(), uRec.contiguous()
        jd = self._prepare(u, uRec)
        if recBarJ is None:
            recBarJ = torch.zeros_like(uRec)
        if method == "gmres":
            return self._gmres(u, uRec, uBarJ, recBarJ, jd, tol, int(maxIter), int(restart), lam0, mu0)
        if method != "fixed-point":
            raise ValueError("method is 'gmres' or 'fixed-point'")
        vfv = self.ev.vfv
        lam = torch.zeros_like(u) if lam0 is None else lam0.clone()
        mu = torch.zeros_like(uRec) if mu0 is None else mu0.clone()
        self.history = []
        r0 = None
        for it in range(1, int(maxIter) + 1):
            uB, rB = rhs_vjp(self.ev, u, uRec, self.dlc, lam)
            rB = rB + recBarJ
            rS, uS = vr_vjp(vfv, mu, 0, -1.0, uRec, u)
            muNew = rB + rS
            dmu = float(torch.linalg.vector_norm(muNew - mu))
            mu = muNew
            r = uB + uS + uBarJ
            rn = float(torch.linalg.vector_norm(r))
            if not np.isfinite(rn + dmu):
                raise FloatingPointError(f"steady adjoint diverged at iteration {it}")
            self.history.append(rn)
            r0 = rn if r0 is None else r0
            if rn <= tol * r0 and dmu <= tol * float(torch.linalg.vector_norm(mu)):
                return lam, mu, it
            lam = lam + self._precond(r, u, jd)
        return lam, mu, int(maxIter)

    def _gmres(self, u, uRec, bu, brec, jd, tol, maxIter, m, lam0, mu0):
        """Restarted right-preconditioned GMRES(m) on A^T x = -b; the Krylov basis lives in two
        [m + 1, n] device arrays and is orthogonalised by classical Gram-Schmidt applied twice
        (two matrix-vector products per step)."""
        nU = u.numel()

        def pack(a, b):
            return torch.cat([a.reshape(-1), b.reshape(-1)])

        def unpack(x):
            return x[:nU].view_as(u), x[nU:].view_as(uRec)

        def A(x):
            a, b = self._apply(u, uRec, *unpack(x))
            return pack(a, b)

        def Minv(y):
            ru, rr = unpack(y)
            return pack(-self._precond(ru.contiguous(), u, jd), -rr)

        b = -pack(bu, brec)
        bnorm = float(torch.linalg.vector_norm(b))
        x = torch.zeros_like(b) if lam0 is None else pack(lam0, mu0 if mu0 is not None else torch.zeros_like(uRec))
        self.history = []
        it = 0
        if bnorm == 0.0:
            return (*[t.clone() for t in unpack(x)], 0)
        V = torch.empty((m + 1, b.numel()), dtype=b.dtype, device=b.device)
        Z = torch.empty((m, b.numel()), dtype=b.dtype, device=b.device)
        while it < maxIter:
            r = b - A(x)
            beta = float(torch.linalg.vector_norm(r))
            self.history.append(beta)
            if not np.isfinite(beta):
                raise FloatingPointError("steady adjoint GMRES diverged")
            if beta <= tol * bnorm:
                break
            V[0] = r / beta
            H = np.zeros((m + 1, m))
            cs, sn = np.zeros(m), np.zeros(m)
            g = np.zeros(m + 1)
            g[0] = beta
            k = 0
            for j in range(m):
                Z[j] = Minv(V[j])
                w = A(Z[j])
                h1 = V[: j + 1] @ w
                w = w - V[: j + 1].T @ h1
                h2 = V[: j + 1] @ w
                w = w - V[: j + 1].T @ h2
                H[: j + 1, j] = (h1 + h2).cpu().numpy()
                H[j + 1, j] = float(torch.linalg.vector_norm(w))
                for i in range(j):  # earlier Givens rotations
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                d = np.hypot(H[j, j], H[j + 1, j])
                cs[j], sn[j] = (1.0, 0.0) if d == 0 else (H[j, j] / d, H[j + 1, j] / d)
                hj1 = H[j + 1, j]
                H[j, j], H[j + 1, j] = cs[j] * H[j, j] + sn[j] * hj1, 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                k, it = j + 1, it + 1
                self.history.append(abs(g[j + 1]))
                if abs(g[j + 1]) <= tol * bnorm or it >= maxIter or hj1 == 0.0:
                    break
                V[j + 1] = w / hj1
            y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
            x = x + torch.as_tensor(y, dtype=b.dtype, device=b.device) @ Z[:k]
        lam, mu = unpack(x)
        return lam.clone(), mu.clone(), it

    def lagrangian(self, u, uRec, lam, mu, J, ev=None):
        """J(u, uRec) + lam . R(u, uRec) + mu . S(uRec, u) on the evaluator ev (default: the
        adjoint's, with its current boundary values); J is a callable
        (ev, u, uRec, deltaLambdaCell) -> float.  Its derivative in a boundary value or a mesh
        parameter at fixed (u, uRec, lam, mu) is dJ/dp of the steady state."""
        ev = self.ev if ev is None else ev
        R = _eval_rhs(ev, u, uRec, self.dlc)
        S = _vr_sweep(ev.vfv, uRec, u, -1.0)
        return float(J(ev, u, uRec, self.dlc)) + float((lam * R).sum()) + float((mu * S).sum())

    def bc_value_gradient(self, u, uRec, lam, mu, J, zones, h=1e-6):
        """dJ/d(bcValues[z, k]) for z in zones, k = 0..4, by central differences of the
        Lagrangian in the boundary value at fixed (u, uRec, lam, mu); [len(zones), 5]."""
        ev = self.ev
        base = np.array(ev._bc[1], copy=True)
        g = np.zeros((len(zones), 5))
        try:
            for iz, z in enumerate(zones):
                for k in range(5):
                    step = h * max(1.0, abs(base[z, k]))
                    vals = []
                    for sgn in (1.0, -1.0):
                        p = base.copy()
                        p[z, k] += sgn * step
                        ev.SetBCValues(p)
                        vals.append(self.lagrangian(u, uRec, lam, mu, J))
                    g[iz, k] = (vals[0] - vals[1]) / (2 * step)
        finally:
            ev.SetBCValues(base)
        return g
