"""Boundary conditions of the Euler path as data.

Codes follow EulerBCType (src/Euler/EulerBC.hpp:35-48).  A mesh carries one zone index
per boundary face (``Mesh.faceZone``); the evaluator maps each zone to a type and a
5-vector of conservative variables (the far-field state of BCFar / BCOutP, the inflow
state of BCIn; unused by the other types), as the reference's BoundaryHandler does.
"""
from __future__ import annotations

import numpy as np

BC_FAR, BC_WALL, BC_WALL_INVIS, BC_WALL_ISOTHERMAL = 1, 2, 3, 4
BC_OUT, BC_OUTP, BC_IN, BC_IN_PSTS, BC_SYM, BC_SPECIAL = 5, 6, 7, 8, 9, 10

NAMES = {"BCFar": BC_FAR, "BCWall": BC_WALL, "BCWallInvis": BC_WALL_INVIS, "BCWallIsothermal": BC_WALL_ISOTHERMAL,
         "BCOut": BC_OUT, "BCOutP": BC_OUTP, "BCIn": BC_IN, "BCInPsTs": BC_IN_PSTS, "BCSym": BC_SYM}
SUPPORTED = (BC_FAR, BC_WALL, BC_WALL_INVIS, BC_WALL_ISOTHERMAL, BC_OUT, BC_OUTP, BC_IN, BC_IN_PSTS, BC_SYM)


def bc_code(t) -> int:
    c = NAMES[t] if isinstance(t, str) else int(t)
    if c not in SUPPORTED:
        raise ValueError(f"boundary type {t!r} is not supported on this path")
    return c


def euler_dir_weights(types, maxOrder: int) -> np.ndarray:
    """[nZone, maxOrder+1] direction weights of boundary faces for the VR functional,
    as EulerEvaluator's constructor sets them (EulerEvaluator.hpp:138-152): BCOut /
    BCSpecial 0 for every order; every other type Dirichlet-like, 1 for the value
    (order 0) and 0 for derivatives."""
    w = np.zeros((len(types), maxOrder + 1))
    for z, t in enumerate(types):
        c = bc_code(t)
        if c in (BC_OUT, BC_SPECIAL):
            continue
        w[z, 0] = 1.0
    return w
