"""Pins of the oracle's in-place SOR reconstruction sweep (SORInstead = true,
src/CFV/VariationalReconstruction_Reconstruction.hxx:534-589).

* Gauss-Seidel identity: cell i of one SOR sweep equals cell i of a Jacobi sweep (the
  separately pinned orc_vr_iter) applied to the mixed state in which the cells scanned
  before i already hold their new values; checked for a shuffled scan order, relaxed and
  unrelaxed, periodic meshes and data-valued boundaries (whose term does not read uRec).
* Euler boundaries read the cell's partially accumulated value (the reference hands
  GetBoundaryRHS the array being updated): with relax = 1 a cell whose only boundary face
  comes first in its face list sees a zero own block there, so it equals the Jacobi
  formula with its own block zeroed (and differs from the Jacobi formula on its old block).
* Fixed point: with boundary data of a polynomial of degree <= p, repeated SOR sweeps
  reach the exact polynomial, the same fixed point as the Jacobi sweep.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import random_rec, random_state, small_mesh, vr_settings


def _oracle(m, order, relax=1.0, **kw):
    import pyoracle as po
    s = vr_settings(order)
    s["jacobiRelax"] = relax
    return po.OracleVR(m, **s, **kw)


def _gauss_seidel_check(o, u, rec, order, cells):
    new = o.vr_sor(rec, u, order)
    pos = np.empty_like(order)
    pos[order] = np.arange(order.size)
    for i in cells:
        mixed = rec.copy()
        before = order[: pos[i]]
        mixed[before] = new[before]
        jac = o.vr_iter(mixed, u)
        assert np.abs(jac[i] - new[i]).max() <= 1e-13 * max(1.0, np.abs(new[i]).max())
    return new


@pytest.mark.parametrize("kind,p,relax", [("2d", 1, 1.0), ("2d", 3, 0.8), ("3d", 2, 1.0), ("3d", 1, 0.7)])
def test_sor_is_gauss_seidel_periodic(kind, p, relax):
    m = small_mesh(kind)
    o = _oracle(m, p, relax)
    N = o.nCell
    u, rec = random_state(N, seed=3), random_rec(N, o.M, 5, seed=4)
    order = np.random.default_rng(5).permutation(N).astype(np.int64)
    new = _gauss_seidel_check(o, u, rec, order, order[[0, 1, N // 3, N // 2, N - 2, N - 1]])
    # natural order by default
    assert np.array_equal(o.vr_sor(rec, u), o.vr_sor(rec, u, np.arange(N, dtype=np.int64)))
    assert np.abs(new - o.vr_iter(rec, u)).max() > 1e-6   # differs from a Jacobi sweep


def test_sor_is_gauss_seidel_with_boundary_data():
    from dndsr_amd import mesh as mm
    m = mm.box_quad_grid(7, 6, 3.5, 3.0, zones=dict(xmin=0, xmax=0, ymin=0, ymax=0), jitter=0.12, seed=2)
    o = _oracle(m, 2, 0.9)
    N = o.nCell
    rng = np.random.default_rng(8)
    o.set_vr_dirichlet(rng.uniform(-1, 1, (o.nFace, o.nGFace, 1)))
    u, rec = rng.uniform(-1, 1, (N, 1)), random_rec(N, o.M, 1, seed=9)
    order = rng.permutation(N).astype(np.int64)
    f2c = m.face2cell.reshape(-1, 2)
    bcells = np.unique(f2c[f2c[:, 1] < 0, 0])
    _gauss_seidel_check(o, u, rec, order, np.concatenate([bcells[:6], order[[0, N - 1]]]))


def test_sor_euler_boundary_reads_partial_value():
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    m = mm.box_quad_grid(6, 5, 3.0, 2.5, zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=0.1, seed=4)
    types = [bc.BC_FAR, bc.BC_OUT, bc.BC_WALL, bc.BC_SYM]
    vals = np.tile([1.0, 0.3, 0.05, 0.0, 1 / 0.56 + 0.5 * 0.0925], (4, 1))
    o = _oracle(m, 1, 1.0, zoneDirWeight=bc.euler_dir_weights(types, 1))
    o.set_bc(types, vals, 1.4)
    N = o.nCell
    u, rec = random_state(N, seed=6), random_rec(N, o.M, 5, seed=7, amp=5e-2)
    order = np.arange(N, dtype=np.int64)
    new = o.vr_sor(rec, u, order)
    f2c = m.face2cell.reshape(-1, 2)
    c2fp, c2f = m.cell2facePtr, m.cell2face
    first = [i for i in range(N)
             if f2c[c2f[c2fp[i]], 1] < 0 and (f2c[c2f[c2fp[i] + 1:c2fp[i + 1]], 1] >= 0).all()]
    assert len(first) >= 3
    for i in first:
        mixed = rec.copy()
        mixed[:i] = new[:i]
        z = mixed.copy()
        z[i] = 0.0
        assert np.abs(o.vr_iter(z, u)[i] - new[i]).max() <= 1e-13
        assert np.abs(o.vr_iter(mixed, u)[i] - new[i]).max() > 1e-8


@pytest.mark.parametrize("kind,fid,p", [("2d", 4, 2), ("3d", 1, 1)])
def test_sor_fixed_point_is_the_exact_polynomial(kind, fid, p):
    from dndsr_amd import bc
    from test_oracle_bc import FIELDS, _box, _dirichlet_values
    f, prm, _ = FIELDS[fid]
    # value-only boundary weights: the boundary data carry no derivatives
    o = _oracle(_box(kind), p, 1.0, zoneDirWeight=bc.euler_dir_weights([bc.BC_FAR], p))
    o.set_vr_dirichlet(_dirichlet_values(o, f))
    u = o.cell_average(fid, prm, 1)
    rec = np.zeros((o.nCell, 1, o.M))
    for _ in range(2000):
        new = o.vr_sor(rec, u)
        done = np.abs(new - rec).max() < 1e-15
        rec = new
        if done:
            break
    assert o.rec_error_l1(fid, prm, u, rec) <= 1e-10
    assert np.abs(o.vr_iter(rec, u) - rec).max() <= 1e-12
