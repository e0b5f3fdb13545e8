"""Pins the oracle's ModelEvaluator::EvaluateRHS (src/CFV/ModelEvaluator.cpp:6-115), the
reference's scalar advection-diffusion test physics, against closed forms (CPU only):

* a constant field has zero residual (test/CFV/test_vr_correctness.py:342-371);
* first order on a uniform grid the residual is the upwind difference -a (u_i - u_{i-1}) / h
  plus sigma (u_{i+1} - 2 u_i + u_{i-1}) / (2 h^2) (the GRP jump term over
  distGRP = 2 min(V) / A = 2 h);
* with the converged reconstruction of sin(kx) cos(ky) the advective residual converges to the
  cell average of -a . grad u at the order of the reconstruction (error ratios under refinement);
* conservation: sum V rhs = 0 on periodic meshes."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from dndsr_amd import mesh as mm


def test_constant_field_has_zero_residual():
    m = mm.periodic_quad_grid(7, 6, 3.0, 3.0, jitter=0.15, seed=2)
    o = po.OracleVR(m, maxOrder=2, intOrder=5)
    u = np.full((o.nCell, 2), 5.0)
    rec = np.zeros((o.nCell, 2, o.M))
    for _ in range(20):
        rec = o.vr_iter(rec, u)
    assert np.abs(rec).max() < 1e-12
    r = o.model_rhs(u, rec, 1.0, 1.0, 0.3)
    assert np.abs(r).max() < 1e-12


@pytest.mark.parametrize("ax,ay,sigma", [(1.0, 0.0, 0.0), (-0.7, 0.0, 0.0), (1.0, 0.0, 0.25), (0.0, 0.0, 0.4)])
def test_first_order_uniform_grid_closed_form(ax, ay, sigma):
    n, L = 16, 4.0
    h = L / n
    m = mm.periodic_quad_grid(n, 3, L, 3 * h)
    o = po.OracleVR(m, maxOrder=1, intOrder=3)
    bary = o.cell_metrics()["bary"]
    u = np.sin(2 * np.pi * bary[:, 0] / L)[:, None]
    r = o.model_rhs(u, None, ax, ay, sigma)
    # neighbours along x in the periodic lattice (cells numbered x fastest)
    ix = np.arange(o.nCell) % n
    left = np.arange(o.nCell) - ix + (ix - 1) % n
    right = np.arange(o.nCell) - ix + (ix + 1) % n
    ui = u[:, 0]
    up = np.where(ax > 0, ui - ui[left], ui[right] - ui)
    ref = -ax * up / h + sigma * (ui[right] - 2 * ui + ui[left]) / (2 * h * h)
    np.testing.assert_allclose(r[:, 0], ref, rtol=1e-12, atol=1e-12)


def _converged_rec(o, u, n=300):
    rec = np.zeros((o.nCell, u.shape[1], o.M))
    for _ in range(n):
        rec = o.vr_iter(rec, u)
    return rec


@pytest.mark.parametrize("order", [1, 2])
def test_advection_residual_converges(order):
    """-a.grad u for u = sin(kx) sin(ky) on [0, L]^2: the error of the residual against the
    exact cell average falls at least at the rate of order p under refinement (p = 1: one,
    p = 2: two)."""
    L, a = 2.0, (1.0, 0.5)
    k = 2 * np.pi / L
    errs = []
    for n in (8, 16, 32):
        m = mm.periodic_quad_grid(n, n, L, L)
        o = po.OracleVR(m, maxOrder=order, intOrder=5)
        u = o.cell_average(2, [L, 2], 1)      # sin(kx) cos(ky)
        rec = _converged_rec(o, u)
        r = o.model_rhs(u, rec, a[0], a[1], 0.0)
        # exact: -(a0 k cos(kx) cos(ky) - a1 k sin(kx) sin(ky)), cell-averaged by the quadrature
        pts, wj = o.cell_quad()
        x, y = pts[..., 0], pts[..., 1]
        ex = -(a[0] * k * np.cos(k * x) * np.cos(k * y) - a[1] * k * np.sin(k * x) * np.sin(k * y))
        ex = (ex * wj).sum(1) / o.cell_metrics()["vol"]
        errs.append(np.abs(r[:, 0] - ex).mean())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates.min() > order - 0.2, (errs, rates)


def test_conservation_periodic_jittered():
    m = mm.periodic_quad_grid(9, 8, 2.0, 2.0, jitter=0.2, seed=7)
    o = po.OracleVR(m, maxOrder=3, intOrder=5)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, (o.nCell, 3))
    rec = rng.uniform(-0.1, 0.1, (o.nCell, 3, o.M))
    r = o.model_rhs(u, rec, 0.8, -0.3, 0.2)
    vol = o.cell_metrics()["vol"]
    tot = (r * vol[:, None]).sum(0)
    assert np.abs(tot).max() <= 1e-13 * np.abs(r * vol[:, None]).sum()
