"""Pins of the oracle's second-order substitute of the VR sweep (maxOrder 1, subs2ndOrder
1 = Gauss rule, 2 = least squares; src/CFV/VariationalReconstruction_Reconstruction.hxx:184-483).

Cell averages of a linear field are its values at the barycentres, so
* least squares recovers the exact gradient in every interior cell of a jittered mesh;
* the Gauss rule (both face interpolations) recovers it on a uniform grid;
* at a flat symmetry plane the ghost of a field odd about the plane (normal momentum
  c * distance, other components constant) is the field's value at the mirrored
  barycentre, so both methods stay exact there too (density and momenta);
* the reconstruction's linear modes are the gradient times the cell's length scale
  (the inverse of the first derivatives of the scaled Taylor basis);
* a uniform state next to a no-slip wall: the Gauss rule's one boundary term gives the
  momentum gradient -rho v A n / V in closed form.
orc_vr_iter and orc_vr_sor dispatch to the substitute; its Jacobian action is zero.
"""
from __future__ import annotations

import numpy as np
import pytest

GAMMA = 1.4


def _oracle(m, method, scheme=1, types=None, vals=None):
    import pyoracle as po
    from dndsr_amd import bc
    kw = {}
    if types is not None:
        kw["zoneDirWeight"] = bc.euler_dir_weights(types, 1)
    o = po.OracleVR(m, maxOrder=1, intOrder=3, subs2ndOrder=method, subs2ndOrderGGScheme=scheme, **kw)
    if types is not None:
        o.set_bc(types, vals, GAMMA)
    return o


def _mesh(dim, jitter):
    from dndsr_amd import mesh as mm
    if dim == 2:
        return mm.box_quad_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=0, ymin=1, ymax=0), jitter=jitter, seed=3)
    return mm.box_hex_grid(5, 4, 4, 1.0, 0.8, 0.8, zones=dict(xmin=0, xmax=0, ymin=0, ymax=0, zmin=1, zmax=0),
                           jitter=jitter, seed=4)


def _odd_field(bary, dim):
    """rho, rho u_t constant, rho u_n = c * (distance to the symmetry plane), rho E constant;
    the plane is y = 0 in 2-D and z = 0 in 3-D; returns (u, exact gradient [5, dim])."""
    n = dim - 1
    N = bary.shape[0]
    u = np.tile([1.0, 0.3, 0.1, 0.05, 2.5], (N, 1))
    u[:, 1 + n] = 0.2 * bary[:, n]
    G = np.zeros((5, dim))
    G[1 + n, n] = 0.2
    return u, G


def _cells(m, zoneOK):
    f2c = m.face2cell.reshape(-1, 2)
    ok = []
    for i in range(m.nCell):
        fs = m.cell2face[m.cell2facePtr[i]:m.cell2facePtr[i + 1]]
        bnd = fs[f2c[fs, 1] < 0]
        if all(zoneOK(m.faceZone[f]) for f in bnd):
            ok.append(i)
    return np.array(ok)


@pytest.mark.parametrize("dim,method,scheme,jitter", [(2, 2, 1, 0.15), (3, 2, 1, 0.1), (2, 1, 1, 0.0),
                                                      (2, 1, 0, 0.0), (3, 1, 1, 0.0), (3, 1, 0, 0.0)])
def test_linear_fields_are_exact(dim, method, scheme, jitter):
    from dndsr_amd import bc
    m = _mesh(dim, jitter)
    far = np.array([1.0, 0.3, 0.1, 0.05, 2.5])
    o = _oracle(m, method, scheme, [bc.BC_FAR, bc.BC_SYM], np.stack([far, far]))
    cm = o.cell_metrics()
    # a general linear field in the interior cells
    rng = np.random.default_rng(1)
    a, G = far + rng.uniform(-0.05, 0.05, 5), rng.uniform(-0.05, 0.05, (5, dim))
    u = a + cm["bary"][:, :dim] @ G.T
    rec = o.vr_2nd(u)
    grad = rec / cm["scale"][:, None, None]
    inner = _cells(m, lambda z: False)
    assert inner.size > 4
    assert np.abs(grad[inner] - G).max() <= 1e-12
    # the odd field with the symmetry plane: also the cells along the plane
    u, G = _odd_field(cm["bary"], dim)
    grad = o.vr_2nd(u) / cm["scale"][:, None, None]
    along = _cells(m, lambda z: z == 1)
    assert along.size > inner.size
    assert np.abs(grad[along][:, :4] - G[:4]).max() <= 1e-12
    # dispatch: the VR sweep, the SOR sweep and the Jacobian action
    assert np.array_equal(o.vr_iter(np.zeros_like(rec), u), o.vr_2nd(u))
    assert np.array_equal(o.vr_sor(np.ones_like(rec), u), o.vr_2nd(u))
    assert not o.vr_iter_diff(np.ones_like(rec), rec, u).any()


def test_gauss_rule_is_not_exact_on_jittered_cells():
    """Sensitivity: the Gauss rule has the face-interpolation error on skewed cells."""
    from dndsr_amd import bc
    m = _mesh(2, 0.15)
    far = np.array([1.0, 0.3, 0.1, 0.05, 2.5])
    o = _oracle(m, 1, 1, [bc.BC_FAR, bc.BC_SYM], np.stack([far, far]))
    cm = o.cell_metrics()
    G = np.zeros((5, 2))
    G[0, 0] = 0.03
    grad = o.vr_2nd(far + cm["bary"][:, :2] @ G.T) / cm["scale"][:, None, None]
    inner = _cells(m, lambda z: False)
    assert np.abs(grad[inner] - G).max() > 1e-4


def test_gauss_rule_wall_term_closed_form():
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    m = mm.box_quad_grid(4, 3, 2.0, 1.5, zones=dict(xmin=0, xmax=0, ymin=1, ymax=0), jitter=0.1, seed=5)
    U = np.array([1.2, 1.2 * 0.3, 1.2 * 0.1, 0.0, 2.5])
    o = _oracle(m, 1, 1, [bc.BC_OUT, bc.BC_WALL], np.stack([U, U]))
    cm, fm = o.cell_metrics(), o.face_metrics()
    grad = o.vr_2nd(np.tile(U, (m.nCell, 1))) / cm["scale"][:, None, None]
    f2c = m.face2cell.reshape(-1, 2)
    for i in _cells(m, lambda z: z == 1):
        fs = m.cell2face[m.cell2facePtr[i]:m.cell2facePtr[i + 1]]
        wall = [f for f in fs if f2c[f, 1] < 0]
        if not wall:
            continue
        f = wall[0]
        A, n, V = fm["area"][f], fm["meanNorm"][f, :2], cm["vol"][i]
        # no-slip ghost: momentum reversed, so (uB - u) / 2 = -rho v
        expect = -np.outer(U[1:3], n) * A / V
        assert np.abs(grad[i, 1:3] - expect).max() <= 1e-13
        assert np.abs(grad[i, [0, 4]]).max() <= 1e-14
