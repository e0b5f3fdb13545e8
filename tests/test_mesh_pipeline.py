"""Mesh pipeline on the host: the Delaunay bump box of the laminar configuration and the
locality renumbering of an unstructured mesh (the role of UnstructuredMesh::ReorderLocalCells,
src/Geom/Mesh/Mesh_Reorder.cpp:662-700).  The renumbering must be a permutation that leaves
every cell-wise result unchanged up to the renumbering (checked through the oracle's residual),
and must shrink the index distance between face neighbours."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import bc
from dndsr_amd import mesh as mm
from helpers import RS, random_rec, random_state, rel_l2


def test_delaunay_box_is_a_conforming_tetrahedral_mesh():
    m = mm.bump_box_delaunay(6)
    assert m.nCell > 6 * 6 ** 3 * 0.9
    P = m.coords[m.cell2node.reshape(-1, 4)]
    vol = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0]) / 6
    assert np.all(vol > 0)
    # the volumes fill the bumped box: int_0^1 (1 - 0.1 sin^2(pi x)) dx = 0.95
    assert abs(vol.sum() - 0.95) < 5e-3
    # every interior face has two cells, every boundary face a zone on its side of the box
    bnd = m.face2cell[:, 1] < 0
    assert np.all(m.faceZone[bnd] >= 0) and np.all(m.faceZone[~bnd] < 0)
    X = m.coords[m.face2node.reshape(-1, 3)[bnd]]
    z = m.faceZone[bnd]
    assert np.allclose(X[z == 0][:, :, 0], 0) and np.allclose(X[z == 1][:, :, 0], 1)
    assert np.allclose(X[z == 3][:, :, 1], 1)
    assert np.all(np.isin(np.round(X[z == 4][:, :, 2], 12), [0.0, 1.0]))
    # a closed surface: the outward face normals of the boundary sum to zero
    e1, e2 = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
    assert np.abs(np.cross(e1, e2).sum(0)).max() < 1e-12


@pytest.mark.parametrize("method", ["rcm", "morton"])
def test_reorder_is_a_renumbering(method):
    m = mm.bump_box_delaunay(6)
    r = mm.reorder_cells(m, method)
    assert mm.bandwidth(r) < mm.bandwidth(m)
    # same cells (as node sets) in a new order, same faces
    c0 = np.sort(m.cell2node.reshape(-1, 4), 1)
    c1 = np.sort(r.cell2node.reshape(-1, 4), 1)
    k0 = {tuple(c): i for i, c in enumerate(c0)}
    perm = np.array([k0[tuple(c)] for c in c1])
    assert np.array_equal(np.sort(perm), np.arange(m.nCell))
    assert r.nFace == m.nFace and np.array_equal(np.bincount(r.faceZone[r.faceZone >= 0]),
                                                 np.bincount(m.faceZone[m.faceZone >= 0]))
    # the residual of the renumbered mesh is the renumbered residual
    types = [bc.BC_IN, bc.BC_IN, bc.BC_WALL, bc.BC_IN, bc.BC_SYM]
    far = random_state(1, seed=3)[0]
    res = []
    for mesh, p in ((m, None), (r, perm)):
        o = po.OracleVR(mesh, maxOrder=1, intOrder=3, zoneDirWeight=bc.euler_dir_weights(types, 1))
        o.set_bc(types, np.stack([far] * 5), 1.4)
        es = o.euler_settings(1.4, RS["Roe"])
        u = random_state(m.nCell, seed=4)
        rec = random_rec(m.nCell, o.M, seed=5)
        if p is not None:
            u, rec = u[p], rec[p]
        rhs = o.eval_rhs(es, u, o.vr_iter(rec, u))
        res.append(rhs if p is None else rhs[np.argsort(p)])
    assert rel_l2(res[1], res[0]) < 1e-13
