"""VariationalReconstruction on the device, with the reference's method names.

Mirrors ``DNDS::CFV::VariationalReconstruction<dim>`` (src/CFV/VariationalReconstruction.hpp:108)
and its Python binding ``CFV.VariationalReconstruction_{2,3}``
(src/CFV/VariationalReconstruction_bind.hpp:18-144).  Field arrays are torch
tensors on the GPU in the device layout of include/dndsr_amd.h (cell index
fastest); torch only provides the memory and the stream, every operator is a
HIP kernel of libdndsr_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

RIEMANN_TYPES = {"Roe": 1, "HLLC": 2, "HLLEP": 3, "HLLEP_V1": 21, "Roe_M1": 11, "Roe_M2": 12, "Roe_M3": 13,
                 "Roe_M4": 14, "Roe_M5": 15, "Roe_M6": 16, "Roe_M7": 17, "Roe_M8": 18}
DIR_WEIGHT = {"Factorial": 0, "HQM_OPT": 1}
GEOM_WEIGHT = {"GWNone": 0, "HQM_SD": 1}


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class Context:
    """One device context (C-ABI ``dndsr_ctx``) that follows torch's current stream: every
    access to the handle re-binds the context when the caller has switched streams (e.g. inside
    ``with torch.cuda.stream(s):``), so library kernels and torch operations on the same tensors
    stay ordered without explicit synchronisation."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("dndsr_amd needs a GPU (no CPU fallback)")
        self.device = device
        h = C.c_void_p()
        check(lib().dndsr_create(C.byref(h), device))
        self._h = h
        self._stream = None
        self.torch_device = torch.device("cuda", device)

    @property
    def h(self):
        """The C-ABI handle, bound to torch's current stream of this device."""
        if self._h is None:
            return None
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            check(lib().dndsr_set_stream(self._h, C.c_void_p(s)))
            self._stream = s
        return self._h

    def sync(self):
        check(lib().dndsr_sync(self.h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().dndsr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VariationalReconstruction:
    """``VariationalReconstruction<dim>``: metrics, VR coefficients and the VR iteration on the device."""

    def __init__(self, mesh, settings: dict | None = None, device: int = 0, ctx: Context | None = None):
        s = dict(maxOrder=1, intOrder=1, dirWeightScheme="Factorial", geomWeightScheme="GWNone",
                 geomWeightBias=0.0, geomWeightPower=0.5, scaleMultiplier=1.0, jacobiRelax=1.0,
                 nIterCellSmoothScale=15, bcWeight=1.0, zoneDirWeight=None, SORInstead=False, subs2ndOrder=0,
                 subs2ndOrderGGScheme=1)
        if settings:
            s.update(settings)
            if "functionalSettings" in settings:
                s.update(settings["functionalSettings"])
        if s.get("bcTypes") is not None and s.get("zoneDirWeight") is None:
            from . import bc as _bc
            s["zoneDirWeight"] = _bc.euler_dir_weights(s["bcTypes"], s["maxOrder"])
        if not isinstance(s["dirWeightScheme"], int):
            s["dirWeightScheme"] = DIR_WEIGHT[s["dirWeightScheme"]]
        if not isinstance(s["geomWeightScheme"], int):
            s["geomWeightScheme"] = GEOM_WEIGHT[s["geomWeightScheme"]]
        self.settings = s
        self.mesh = mesh
        self.dim = mesh.dim
        self.ctx = ctx or Context(device)
        self._constructed = False
        self._dirichlet = None

    # ---- construction (ConstructMetrics / ConstructBaseAndWeight / ConstructRecCoeff)
    def ConstructMetrics(self):
        if self._constructed:
            return
        m = self.mesh
        keep = [np.ascontiguousarray(x) for x in (
            m.coords, m.cellType, m.cell2nodePtr, m.cell2node, m.cell2facePtr, m.cell2face, m.faceType,
            m.face2nodePtr, m.face2node, m.face2cell, m.faceShift)]
        fz = getattr(m, "faceZone", None)
        fz = None if fz is None else np.ascontiguousarray(fz, np.int32)
        zw = self.settings.get("zoneDirWeight")
        zw = None if zw is None else np.ascontiguousarray(zw, np.float64)
        keep += [fz, zw]
        md = _lib.MeshDesc(m.dim, m.nNode, keep[0].ctypes.data_as(_lib.dp), m.nCell, keep[1].ctypes.data_as(_lib.i32p),
                           keep[2].ctypes.data_as(_lib.i64p), keep[3].ctypes.data_as(_lib.i64p),
                           keep[4].ctypes.data_as(_lib.i64p), keep[5].ctypes.data_as(_lib.i64p), m.nFace,
                           keep[6].ctypes.data_as(_lib.i32p), keep[7].ctypes.data_as(_lib.i64p),
                           keep[8].ctypes.data_as(_lib.i64p), keep[9].ctypes.data_as(_lib.i64p),
                           keep[10].ctypes.data_as(_lib.dp), None if fz is None else fz.ctypes.data_as(_lib.i32p),
                           int(getattr(m, "nBCZone", 0)))
        s = self.settings
        vs = _lib.VRSettings(s["maxOrder"], s["intOrder"], s["dirWeightScheme"], s["geomWeightScheme"],
                             s["geomWeightBias"], s["geomWeightPower"], s["scaleMultiplier"], s["jacobiRelax"],
                             s["nIterCellSmoothScale"], s["bcWeight"],
                             None if zw is None else zw.ctypes.data_as(_lib.dp), int(bool(s["SORInstead"])),
                             int(s["subs2ndOrder"]), int(s["subs2ndOrderGGScheme"]))
        check(lib().dndsr_vr_construct(self.ctx.h, C.byref(md), C.byref(vs)))
        nC, nF = C.c_int64(), C.c_int64()
        M, K, G, Gc, nCol = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        gv = C.c_double()
        check(lib().dndsr_mesh_info(self.ctx.h, C.byref(nC), C.byref(nF), C.byref(M), C.byref(K), C.byref(G),
                                    C.byref(Gc), C.byref(nCol), C.byref(gv)))
        self.nCell, self.nFace, self.M, self.maxFC = nC.value, nF.value, M.value, K.value
        self.nGFace, self.nGCell, self.nColor, self.globalVol = G.value, Gc.value, nCol.value, gv.value
        self._constructed = True

    def ConstructBaseAndWeight(self, id2faceDircWeight=None):
        """id2faceDircWeight(zone, order) multiplies the weights of boundary faces
        (VariationalReconstruction.cpp:438-441); it must come before the device build."""
        if id2faceDircWeight is not None:
            nz = int(getattr(self.mesh, "nBCZone", 0))
            P = self.settings["maxOrder"]
            w = np.array([[id2faceDircWeight(z, p) for p in range(P + 1)] for z in range(nz)], np.float64)
            if self._constructed:
                old = self.settings.get("zoneDirWeight")
                if old is None or not np.array_equal(np.asarray(old), w):
                    raise RuntimeError("boundary weights must be set before the VR coefficients are built")
            self.settings["zoneDirWeight"] = w
        self.ConstructMetrics()

    def ConstructRecCoeff(self):
        self.ConstructMetrics()

    def construct(self):
        self.ConstructMetrics()
        return self

    # ---- arrays
    def BuildUDof(self, nVars: int = 5, dtype=torch.float64):
        return torch.zeros((nVars, self.nCell), dtype=dtype, device=self.ctx.torch_device)

    def BuildURec(self, nVars: int = 5):
        return torch.zeros((nVars, self.M, self.nCell), dtype=torch.float64, device=self.ctx.torch_device)

    def to_device(self, rows: np.ndarray, out=None):
        """Reference row layout [nCell, nComp] (uRec: [nCell, nVars, M]) -> device tensor."""
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        nComp = int(np.prod(rows.shape[1:]))
        if out is None:
            shape = (rows.shape[1], rows.shape[2], self.nCell) if rows.ndim == 3 else (nComp, self.nCell)
            out = torch.empty(shape, dtype=torch.float64, device=self.ctx.torch_device)
        check(lib().dndsr_h2d_cellfield(self.ctx.h, rows.ctypes.data_as(C.c_void_p), _ptr(out), nComp))
        return out

    def to_host(self, dev) -> np.ndarray:
        """Device tensor -> reference row layout."""
        nComp = int(np.prod(dev.shape[:-1]))
        rows = np.empty((self.nCell, nComp))
        check(lib().dndsr_d2h_cellfield(self.ctx.h, _ptr(dev), rows.ctypes.data_as(C.c_void_p), nComp))
        if dev.dim() == 3:
            rows = rows.reshape(self.nCell, dev.shape[0], dev.shape[1])
        return rows

    # ---- metrics (host copies, reference row layout)
    def cell_metrics(self):
        N, M = self.nCell, self.M
        vol, sc, sm = np.zeros(N), np.zeros(N), np.zeros(N)
        bary, cent, mom = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, M))
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        check(lib().dndsr_get_cell_metrics(self.ctx.h, P(vol), P(bary), P(cent), P(sc), P(sm), P(mom)))
        return dict(vol=vol, bary=bary, cent=cent, scale=sc, smooth=sm, moment=mom)

    def face_metrics(self):
        F, G = self.nFace, self.nGFace
        area, fl = np.zeros(F), np.zeros(F)
        mn, pts, nrm = np.zeros((F, 3)), np.zeros((F, G, 3)), np.zeros((F, G, 3))
        wj, fw = np.zeros((F, G)), np.zeros((F, self.settings["maxOrder"] + 1))
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        check(lib().dndsr_get_face_metrics(self.ctx.h, P(area), P(mn), P(pts), P(nrm), P(wj), P(fl), P(fw)))
        return dict(area=area, meanNorm=mn, pts=pts, norms=nrm, wJDet=wj, faceL=fl, faceWeight=fw)

    def cell_quad(self):
        pts, wj = np.zeros((self.nCell, self.nGCell, 3)), np.zeros((self.nCell, self.nGCell))
        check(lib().dndsr_get_cell_quad(self.ctx.h, pts.ctypes.data_as(C.c_void_p), wj.ctypes.data_as(C.c_void_p)))
        return pts, wj

    def get_matrixAAInvB(self):
        """(matrixAAInvB [N, K+1, M, M], vectorAInvB [N, K, M]) in the reference's block order."""
        N, M, K = self.nCell, self.M, self.maxFC
        A, b = np.zeros((N, K + 1, M, M)), np.zeros((N, K, M))
        check(lib().dndsr_get_vr_coeffs(self.ctx.h, A.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))
        return A, b

    def set_matrixAAInvB(self, A, b):
        A = np.ascontiguousarray(A, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        check(lib().dndsr_set_vr_coeffs(self.ctx.h, A.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))

    def colors(self):
        c = np.zeros(self.nCell, np.int32)
        check(lib().dndsr_get_colors(self.ctx.h, c.ctypes.data_as(C.c_void_p)))
        return c

    def cell_average(self, f, nVars: int):
        """Cell means of a point function f(pts[..., 3]) -> [..., nVars] by cell quadrature
        (EulerEvaluator.hxx:1046-1068)."""
        pts, wj = self.cell_quad()
        vals = f(pts).reshape(self.nCell, self.nGCell, nVars)
        vol = self.cell_metrics()["vol"]
        return np.einsum("ngv,ng->nv", vals, wj) / vol[:, None]

    # ---- reconstruction (VariationalReconstruction_Reconstruction.hxx:485-670)
    def DoReconstructionIter(self, uRec, uRecNew, u, putIntoNew: bool = True, recordInc: bool = False,
                             uRecIsZero: bool = False, relax: float = -1.0):
        """putIntoNew=False leaves the new iterate in uRec (as the reference does).  With
        settings SORInstead (and no recordInc) uRec is swept in place, colour by colour; for
        putIntoNew the contents of uRec and uRecNew are then swapped, as the reference swaps
        the arrays."""
        nVars = u.shape[0]
        mode = 0 if not recordInc else (2 if uRecIsZero else 1)
        if recordInc and not putIntoNew:
            raise ValueError("recordInc needs putIntoNew (the -RHS goes into uRecNew)")
        if self.settings["SORInstead"] and not recordInc:
            self.DoReconstructionSOR(uRec, u, relax)
            if putIntoNew:
                old = uRecNew.clone()
                uRecNew.copy_(uRec)
                uRec.copy_(old)
            return
        check(lib().dndsr_vr_iter(self.ctx.h, nVars, _ptr(uRec), _ptr(uRecNew), _ptr(u), mode, float(relax)))
        if not putIntoNew:
            uRec.copy_(uRecNew)

    def DoReconstructionSOR(self, uRec, u, relax: float = -1.0):
        """One multicolour SOR sweep, uRec updated in place (cells colour by colour, colors())."""
        check(lib().dndsr_vr_iter(self.ctx.h, u.shape[0], _ptr(uRec), None, _ptr(u), 4, float(relax)))

    def DoReconstructionIterDiff(self, uRec, uRecDiff, uRecNew, u=None):
        """uRecNew = uRecDiff - sum_f A^-1 B_f uRecDiff_j - A^-1 BCdiff; uRec and u feed the
        boundary term (None on meshes without boundary faces)."""
        check(lib().dndsr_vr_iter_diff(self.ctx.h, uRecDiff.shape[0], _ptr(uRec), _ptr(u), _ptr(uRecDiff),
                                       _ptr(uRecNew)))

    def set_vr_dirichlet(self, values):
        """Boundary point values for the VR sweep when the Euler boundary table does not apply
        (any nVars): device tensor [nVars, nGFace, nFace] (boundary faces read); kept by reference."""
        if values is None:
            check(lib().dndsr_set_vr_dirichlet(self.ctx.h, 0, None))
            self._dirichlet = None
            return
        self._dirichlet = values.contiguous()
        check(lib().dndsr_set_vr_dirichlet(self.ctx.h, int(values.shape[0]), _ptr(self._dirichlet)))
