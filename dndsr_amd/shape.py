"""Hicks-Henne shape parametrisation of a wall bump, and the shape gradient of a steady
functional through the steady adjoint.  The reference has neither: they make up the 2-D
discrete-adjoint workload (converged Mach 0.7, Re 5e3 laminar flow over a lower-wall bump
given by eight Hicks-Henne functions, mixed triangle/quadrilateral mesh, drag functional).

Hicks-Henne functions (Hicks & Henne, J. Aircraft 15(7), 1978), on the bump chord [0, 1]:

    f_k(x) = sin(pi x^(ln 0.5 / ln x_k))^t,   x_k = k / (n + 1),  k = 1..n,

zero outside the chord; f_k peaks (value 1) at x_k.  The lower wall is y_w(x) =
sum_k alpha_k f_k(x), and every mesh node moves vertically by y_w(x) (1 - y / H), so the
top boundary stays fixed and the mesh is a linear function of alpha.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import mesh as mm


def hicks_henne(x, n: int = 8, t: float = 3.0):
    """[n, len(x)] values of the n Hicks-Henne functions at the chord positions x."""
    x = np.asarray(x, np.float64)
    xk = np.arange(1, n + 1) / (n + 1)
    e = np.log(0.5) / np.log(xk)
    inside = (x > 0) & (x < 1)
    xc = np.where(inside, x, 0.5)
    return np.where(inside[None], np.sin(np.pi * xc[None] ** e[:, None]) ** t, 0.0)


class HicksHenneBumpChannel:
    """Channel [x0, x0 + L] x [0, H] of mixed cells (mesh.mixed_grid: squares in the right half
    of the lattice cut into two triangles), the bump chord [0, 1] on the lower wall.  Zones:
    0 the lower wall (BCWall), 1 inlet and outlet (BCFar), 2 the top (BCSym: the far-field
    ghost changes form where the normal velocity changes sign, so a far-field boundary along
    the stream keeps the residual from converging).  mesh(alpha) keeps the topology and
    numbering and moves the nodes; nCell = 1.5 nx ny."""

    BC_TYPES = ("BCWall", "BCFar", "BCSym")

    def __init__(self, nx: int, ny: int, x0: float = -1.0, L: float = 3.0, H: float = 1.5, n: int = 8,
                 t: float = 3.0):
        self.n, self.t, self.H = int(n), float(t), float(H)
        self.base = mm.mixed_grid(nx, ny, L, H, x0=x0, zones=dict(ymin=0, xmin=1, xmax=1, ymax=2))
        self.X0 = self.base.coords.copy()
        self._f = hicks_henne(self.X0[:, 0], self.n, self.t)  # [n, nNode]

    def wall(self, alpha, x):
        """The lower-wall height y_w(x)."""
        return np.asarray(alpha, np.float64) @ hicks_henne(x, self.n, self.t)

    def coords(self, alpha):
        X = self.X0.copy()
        X[:, 1] += (np.asarray(alpha, np.float64) @ self._f) * (1.0 - self.X0[:, 1] / self.H)
        return X

    def mesh(self, alpha):
        return dataclasses.replace(self.base, coords=self.coords(alpha))

    @staticmethod
    def flow(M: float = 0.7, Re: float = 5e3, gamma: float = 1.4):
        """Free stream (rho 1, speed M, a = 1) and laminar gas (constant mu = M / Re on the
        unit chord, Pr 0.72) of the workload: (bcValues for the three zones, idealGasProperty)."""
        far = np.array([1.0, M, 0.0, 0.0, 1.0 / gamma / (gamma - 1.0) + 0.5 * M * M])
        return np.stack([far] * 3), dict(Rgas=1.0, muGas=M / Re, prGas=0.72, muModel=0, gamma=gamma)


def shape_gradient(adj, u, uRec, lam, mu, J, build, alpha, h: float = 1e-6):
    """dJ/dalpha of the steady state from the multipliers (lam, mu) of adj (a SteadyAdjoint
    solved at the state (u, uRec) for the mesh at alpha): central differences in each alpha_k
    of the Lagrangian J + lam . R + mu . S at fixed (u, uRec, lam, mu), each on an evaluator
    that build(alpha') makes for the moved mesh (same settings and numbering).  The state
    stays fixed, so each parameter costs two metric/VR-matrix set-ups and two residual and VR
    evaluations, whatever the size of the state."""
    alpha = np.asarray(alpha, np.float64)
    g = np.zeros(len(alpha))
    for k in range(len(alpha)):
        vals = []
        for sgn in (1.0, -1.0):
            a = alpha.copy()
            a[k] += sgn * h
            ev = build(a)
            try:
                vals.append(adj.lagrangian(u, uRec, lam, mu, J, ev=ev))
            finally:
                ev.ctx.close()
        g[k] = (vals[0] - vals[1]) / (2 * h)
    return g
