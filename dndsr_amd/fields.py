"""Analytic initial fields evaluated at quadrature points (vectorised numpy).

isentropic_vortex: SpecialFields::IsentropicVortex10 (src/Euler/SpecialFields.hpp:41-75), with
  the background velocity as a parameter; (1, 1) reproduces the reference, and
  (1, 1) * M a_inf / sqrt(2) gives a free stream of Mach M along the diagonal.
taylor_green: specialBuiltinInitializer 3 (src/Euler/EulerEvaluator.hxx:1098-1135).
"""
from __future__ import annotations

import numpy as np


def isentropic_vortex(pts, chi=5.0, gamma=1.4, ub=(1.0, 1.0), center=(5.0, 5.0)):
    x, y = pts[..., 0], pts[..., 1]
    xc, yc = center
    r = np.sqrt((x - xc) ** 2 + (y - yc) ** 2)
    dT = -(gamma - 1) / (8 * gamma * np.pi ** 2) * chi ** 2 * np.exp(1 - r ** 2)
    dux = chi / 2 / np.pi * np.exp((1 - r ** 2) / 2) * -(y - yc)
    duy = chi / 2 / np.pi * np.exp((1 - r ** 2) / 2) * (x - xc)
    T = dT + 1
    ux, uy = dux + ub[0], duy + ub[1]
    rho = T ** (1 / (gamma - 1))
    p = T * rho
    E = 0.5 * (ux ** 2 + uy ** 2) * rho + p / (gamma - 1)
    return np.stack([rho, rho * ux, rho * uy, np.zeros_like(rho), E], axis=-1)


def iv_background_for_mach(mach, gamma=1.4):
    a = np.sqrt(gamma)  # rho = p = T = 1 far from the vortex
    s = mach * a / np.sqrt(2.0)
    return (s, s)


def taylor_green(pts, M0=0.1, gamma=1.4):
    x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
    ux = np.sin(x) * np.cos(y) * np.cos(z)
    uy = -np.cos(x) * np.sin(y) * np.cos(z)
    p = 1.0 / (gamma * M0 ** 2) + 1.0 / 16 * ((np.cos(2 * x) + np.cos(2 * y)) * (2 + np.cos(2 * z)))
    rho = gamma * M0 ** 2 * p
    E = 0.5 * (ux ** 2 + uy ** 2) * rho + p / (gamma - 1)
    return np.stack([rho, rho * ux, rho * uy, np.zeros_like(rho), E], axis=-1)
