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
    """The residual's transpose chains transposed Jacobi sweeps; with SORInstead the solver's
    frhs sweeps in place (multicolour SOR) and with subs2ndOrder it uses the gradient
    substitute, neither of which this chain differentiates."""
    if vfv.settings.get("SORInstead"):
        raise ValueError("the residual adjoint differentiates Jacobi VR sweeps; the settings select SORInstead "
                         "(the transpose of the in-place SOR sweep is not provided)")
    if vfv.settings.get("subs2ndOrder"):
        raise ValueError("the residual adjoint does not cover the second-order substitute (subs2ndOrder)")


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
