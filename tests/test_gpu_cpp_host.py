"""The C++17 host mirror (include/dndsr_amd.hpp), driven by a compiled caller
(tests/cpp/host_mirror_main.cpp), against the CPU oracle: one VR sweep (Jacobi, or SOR
in the device's colour order with SORInstead),
EvaluateDt (local), EvaluateRHS, LUSGSMatrixInit and a forward Gauss-Seidel sweep in
the device's colour order, at relative L2 <= 1e-12."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from helpers import RS, colour_scan_order, random_rec, random_state, rel_l2, small_mesh, vr_settings

pytestmark = pytest.mark.gpu


def _write_inputs(d, m, u, rec, par):
    os.makedirs(d, exist_ok=True)
    arrays = dict(coords=m.coords, cellType=m.cellType, faceType=m.faceType, cell2nodePtr=m.cell2nodePtr,
                  cell2node=m.cell2node, cell2facePtr=m.cell2facePtr, cell2face=m.cell2face,
                  face2nodePtr=m.face2nodePtr, face2node=m.face2node, face2cell=m.face2cell, faceShift=m.faceShift,
                  u=u, uRec=rec, params=np.asarray(par, np.float64))
    for k, v in arrays.items():
        np.save(os.path.join(d, k + ".npy"), np.ascontiguousarray(v))


@pytest.mark.parametrize("kind,order,rs,sor", [("2d", 1, "Roe", 0), ("2d", 2, "HLLEP", 0), ("3d", 2, "HLLEP", 0),
                                                ("2d", 2, "Roe", 1)])
def test_cpp_host_mirror_matches_oracle(require_gpu, tmp_path, kind, order, rs, sor):
    import pyoracle as po
    from dndsr_amd import build as b
    print(f"\n[cpp host] {kind} p{order} {rs} SOR={sor}", flush=True)
    m = small_mesh(kind, jitter=0.1)
    s = vr_settings(order, intOrder=5)
    o = po.OracleVR(m, **s)
    N, M = o.nCell, o.M
    u = random_state(N, 31)
    rec = random_rec(N, M, 5, seed=32)
    CFL = 2.0
    _write_inputs(str(tmp_path / "in"), m, u, rec, [m.dim, order, 5, RS[rs], CFL, sor])
    os.makedirs(tmp_path / "out")
    r = subprocess.run([b.MIRROR_TEST_BIN, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ld = lambda k: np.load(tmp_path / "out" / f"{k}.npy")  # noqa: E731
    es = o.euler_settings(1.4, RS[rs], 0, 1.0, 1.0)
    recO = o.vr_sor(rec, u, colour_scan_order(ld("colors"))) if sor else o.vr_iter(rec, u)
    assert rel_l2(ld("uRec"), recO) <= 1e-12
    dt, lam, dlc, dtMin = o.eval_dt(es, u, CFL, 1e100, True)
    assert rel_l2(ld("dt"), dt) <= 1e-12
    assert abs(ld("dtMin")[0] - dtMin) <= 1e-12 * dtMin
    rhs = o.eval_rhs(es, u, recO, dlc)
    assert rel_l2(ld("rhs"), rhs) <= 1e-12
    np.testing.assert_allclose(ld("norm"), np.abs(rhs).sum(0), rtol=1e-12)
    jd = o.lusgs_init(dt, 1e100, 1.0, lam)
    assert rel_l2(ld("jd"), jd) <= 1e-12
    z = np.zeros_like(u)
    inc, _ = o.update_sgs(es, 1.0, rhs, u, z, z, jd, lam, True, True, False,
                          scanOrder=colour_scan_order(ld("colors")))
    assert rel_l2(ld("inc"), inc) <= 1e-12
