"""dndsr_eval_rhs_noalloc (the EvaluateRHS entry point that allocates nothing) against
dndsr_eval_rhs, through the C-ABI: two calls in a row on one GPU, the second on a smaller mesh
(so the shared static scratch holds the first call's face fluxes when the second starts), each
bit-identical to the allocating entry point on the same inputs; a mesh beyond the scratch's
capacity is refused with an error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import RS, random_rec, random_state

pytestmark = pytest.mark.gpu


def _setup(m, order, rs):
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    g = VariationalReconstruction(m, dict(maxOrder=order, intOrder=order + 3 if order < 3 else 5)).construct()
    ev = EulerEvaluator(g, dict(riemannSolverType=rs))
    u = g.to_device(random_state(g.nCell, 11))
    rec = g.to_device(random_rec(g.nCell, g.M, seed=12))
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, u, rec, CFL=0.5)
    return g, ev, u, rec


def _rhs(fn, g, ev, u, rec):
    from dndsr_amd._lib import check, lib
    from dndsr_amd.cfv import _ptr
    out = g.BuildUDof(5)
    check(getattr(lib(), fn)(ev.ctx.h, C.byref(ev.es), _ptr(u), _ptr(rec), _ptr(ev.deltaLambdaCell), _ptr(out)))
    ev.ctx.sync()
    return out.cpu().numpy()


def test_noalloc_rhs_matches_allocating_entry_point(require_gpu):
    from dndsr_amd import mesh as mm
    cases = [
        ("3-D 20x18x16 hex p2 HLLE+", mm.periodic_hex_grid(20, 18, 16, 1.0, 1.1, 0.9, jitter=0.1, seed=2), 2,
         "HLLEP"),
        ("2-D 9x7 quad p1 Roe", mm.periodic_quad_grid(9, 7, 1.0, 0.8, jitter=0.1, seed=3), 1, "Roe"),
    ]
    for name, m, order, rs in cases:
        print(f"\n[noalloc] {name}", flush=True)
        g, ev, u, rec = _setup(m, order, rs)
        a = _rhs("dndsr_eval_rhs_noalloc", g, ev, u, rec)
        b = _rhs("dndsr_eval_rhs", g, ev, u, rec)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), np.abs(a - b).max()


def test_noalloc_rejects_meshes_beyond_its_scratch(require_gpu):
    from dndsr_amd import mesh as mm
    print("\n[noalloc] 72^3 periodic hex p1 (1119744 faces): refused", flush=True)
    m = mm.periodic_hex_grid(72, 72, 72, 1.0, 1.0, 1.0)
    g, ev, u, rec = _setup(m, 1, "Roe")
    assert g.nFace > 1 << 20
    with pytest.raises(RuntimeError, match="more faces than its scratch"):
        _rhs("dndsr_eval_rhs_noalloc", g, ev, u, rec)
