"""Benchmark of the compact-FV hot path on one MI355X.

Workload (the reference-path configuration the metric is quoted on): 3-D Euler
Taylor-Green vortex, Mach 0.1, periodic [0, 2 pi]^3 with 64^3 = 262144 uniform
hexahedra, 3rd-order VR (p2, 9 modes per variable, intOrder 5), HLLE+ flux, FP64,
explicit SSP-RK3 at CFL 0.5.  One step = EvaluateDt + 3 x (one Jacobi VR sweep +
EvaluateRHS + stage combination), all on the device (dndsr_ssprk3).

value = cell-updates per second per RHS evaluation = nCell * 3 * steps / time.
The roofline entry is the dominant kernel: algorithmic bytes per launch divided by
its average launch time, measured with HIP events on its stream during the timed
region; the same kernels are also timed as back-to-back loops under one event pair,
and the device's own read and copy streams give the measured bandwidth the fractions
are also quoted against.  The second half of the metric, time to a 1e-8 residual with
FP64 and with FP32 sweeps, runs on an inviscid hex bump and on the laminar bump box
(2M Delaunay tetrahedra).  cpu_baseline times the CPU oracle on the 64^3 workload
with all physical host cores and with one core (median of five repetitions), and
checks the device results against it (error_vs_oracle).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E datasheet peak
METRIC = "cell-updates/s per RHS eval; time-to-1e-8 residual, mixed vs FP64 GPU"
HERE = os.path.dirname(os.path.abspath(__file__))


def alg_bytes(dim, M, nV, n_fc, f_c, G_f):
    """Algorithmic HBM bytes per cell and launch: each array the kernels read or write
    counted once at its stored precision, 4-byte indices, basis values recomputed (not
    tabulated); the face-flux intermediate of EvaluateRHS is not counted.
    VR Jacobi sweep: 8 [2 M nV + nV + n_fc (M^2 + M)] + 4 n_fc
        (uRec in/out, u, A^-1 B blocks and A^-1 b per face slot, neighbour index)
    EvaluateRHS:     8 [nV (2 + M) + M + d + 3] + f_c [8 + 8 (G_f (2 d + 1) + 3)]
        per cell: u, rhs out, uRec, moments, centre, 1/scale, volume, deltaLambda;
        per face (f_c faces per cell): the two cell indices, point coordinates, unit
        normals and w JDet at G_f points, the periodic shift"""
    b_vr = 8 * (2 * M * nV + nV + n_fc * (M * M + M)) + 4 * n_fc
    b_rhs = 8 * (nV * (2 + M) + M + dim + 3) + f_c * (8 + 8 * (G_f * (2 * dim + 1) + 3))
    return b_vr, b_rhs


def read_traffic(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE),
    recorded in profiles/traffic_<kernel_tag>.txt as `bytes_per_launch: X`."""
    p = os.path.join(HERE, "profiles", f"traffic_{kernel_tag}.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def host_cpu_info():
    """nproc (the CPUs this process may run on), and lscpu's model, sockets, cores per socket and
    threads per core; physical cores available = min(affinity, sockets x cores per socket)."""
    import subprocess
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    info = {"nproc": aff}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {ln.split(":", 1)[0].strip(): ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ":" in ln}
        info.update({"model": kv.get("Model name"), "sockets": int(kv.get("Socket(s)", "1")),
                     "cores_per_socket": int(kv.get("Core(s) per socket", "0") or 0),
                     "threads_per_core": int(kv.get("Thread(s) per core", "1") or 1)})
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    phys = info.get("sockets", 1) * info.get("cores_per_socket", 0)
    tpc = info.get("threads_per_core", 1)
    info["physical_cores_available"] = min(aff, phys) if phys else max(1, aff // tpc)
    return info


def cpu_baseline(n=64, reps=5, evals_all=4, evals_one=1, checks=None):
    """The reference-path CPU baseline: the oracle's frhs (one Jacobi VR sweep + EvaluateRHS, the
    OpenMP face-buffer + cell-gather form of EvaluateRHS.hxx:153-159,640-662) on the benchmark
    configuration itself (TGV n^3 hex, p2, HLLE+, FP64), timed with all physical cores available
    and with one core: after one warm-up evaluation, the median of `reps` repetitions of
    `evals_all` (all cores) or `evals_one` (one core) evaluations; value = cells / median seconds
    per evaluation.  The oracle's setup (metrics and VR coefficients) is not timed.
    checks: device results to check against the oracle on the same inputs (relative L2 of the
    difference, returned as error_vs_oracle): "tgv" the frhs of the 64^3 state (VR sweep from zero
    and EvaluateRHS), "laminar" the state after two implicit iterations on the tetrahedral box,
    "bump" the same on the hex bump."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle as po
    from dndsr_amd import mesh as mm
    info = host_cpu_info()
    cores = int(info["physical_cores_available"])
    t_setup = time.perf_counter()
    po.set_threads(cores, rhsGather=True)
    m = mm.periodic_hex_grid(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    o = po.OracleVR(m, maxOrder=2, intOrder=5)
    u = o.cell_average(11, [0.1, 1.4], 5)
    es = o.euler_settings(1.4, 3, 0, 1.0, 1.0)
    _, _, dlc, _ = o.eval_dt(es, u, 0.5, 1e100, False)
    rec = o.vr_iter(np.zeros((o.nCell, 5, o.M)), u)
    t_setup = time.perf_counter() - t_setup

    def frhs():
        nonlocal rec
        rec = o.vr_iter(rec, u)
        o.eval_rhs(es, u, rec, dlc)

    res = {}
    for label, threads, k in (("all", cores, evals_all), ("one", 1, evals_one)):
        po.set_threads(threads, rhsGather=True)
        frhs()  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(k):
                frhs()
            times.append((time.perf_counter() - t0) / k)
        res[label] = (float(np.median(times)), times, k)
    v_all = o.nCell / res["all"][0]
    v_one = o.nCell / res["one"][0]
    out = {"value": v_all, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "value_1core": v_one, "cores_used_1core": 1,
           "sample": (f"the benchmark workload itself (TGV {n}^3 = {o.nCell} hex, p2, HLLE+, FP64): oracle frhs = "
                      f"Jacobi VR sweep + EvaluateRHS (OpenMP face buffer + cell gather); median of {reps} "
                      f"repetitions of {evals_all} evaluations on {cores} cores "
                      f"({res['all'][0] * 1e3:.1f} ms each) and of {evals_one} on 1 core "
                      f"({res['one'][0] * 1e3:.1f} ms); oracle setup {t_setup:.1f} s untimed"),
           "host": info}
    errors = {}
    if checks:
        po.set_threads(cores)
        rl2 = lambda a, b: float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))  # noqa: E731
        c = checks.get("tgv")
        if c is not None:
            u0 = c["u0"]
            _, _, dlc0, _ = o.eval_dt(es, u0, 0.5, 1e100, False)
            rec0 = o.vr_iter(np.zeros((o.nCell, 5, o.M)), u0)
            errors["tgv64_frhs"] = {"rel_l2": rl2(c["rhs"], o.eval_rhs(es, u0, rec0, dlc0)),
                                    "what": "one Jacobi VR sweep from zero + EvaluateRHS on the initial state"}
        del o
        for key in ("laminar", "bump"):
            c = checks.get(key)
            if c is None:
                continue
            t0 = time.perf_counter()
            ol = po.OracleVR(c["mesh"], maxOrder=1, intOrder=3, zoneDirWeight=c["zoneDirWeight"])
            ol.set_bc(c["types"], c["values"], 1.4)
            el = po.OracleVR.euler_settings(1.4, 1, 0, 1.0, 1.0, viscous=int(c["viscous"]), **c["gas"])
            uo, _, nIt, _, _, _ = ol.implicit_euler_step(el, c["u0"], np.zeros((ol.nCell, 5, ol.M)), 1e100,
                                                          CFL=c["CFL"], sgsIter=c["sgsIter"], maxIter=c["iters"],
                                                          minIter=c["iters"], tol=0.0, scanOrder=c["scanOrder"])
            errors[key] = {"rel_l2": rl2(c["u"], uo), "what": f"state after {nIt} implicit iterations (FP64 sweeps) "
                           f"from the free stream, oracle in the device's colour order",
                           "oracle_s": round(time.perf_counter() - t0, 1)}
            del ol
    out["error_vs_oracle"] = errors
    return out


def implicit_bump_time_to_tol(tol=1e-8, CFL=1000.0, sgsIter=3, max_iter=6000, shape=(160, 48, 32)):
    """Second half of the metric: wall time for the implicit LU-SGS pseudo-time loop
    (ImplicitEulerDualTimeStep::Step with an infinite physical step, i.e. a steady solve)
    to drive the density residual, relative to its running maximum, below tol -- with FP64
    sweeps and with FP32 sweeps (residual, reconstruction and update in FP64).
    Case: steady inviscid Mach 0.5 flow over a sinusoidal wall bump (height 0.1 of the
    channel height) on a 3-D hex mesh, p1 VR (2nd order), Roe flux; slip wall below, the
    free-stream state imposed through the Riemann flux at inlet, outlet and top (BCIn),
    spanwise periodic.  Reports iterations, time and the relative L2 difference of the
    two converged states."""
    import torch
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    gamma = 1.4
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / gamma / (gamma - 1) + 0.5 * 0.25])  # Mach 0.5
    types = [bc.BC_IN, bc.BC_IN, bc.BC_WALL_INVIS, bc.BC_IN]
    m = mm.bump_channel_hex(*shape, jitter=0.1, seed=1)
    vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    m = mm.colour_sorted(m, VariationalReconstruction(m, vs).construct().colors())  # coalesced colour passes
    g = VariationalReconstruction(m, vs).construct()
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * len(types))))
    u0 = g.to_device(np.tile(far, (g.nCell, 1)))
    out, states = {}, {}
    for name, fp32 in (("fp64", False), ("mixed", True)):
        best = None
        for rep in range(2):  # first run warms allocations and kernels
            u = u0.clone()
            rec = g.BuildURec(5)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            n, res, rel, hist = ev.ImplicitEulerStep(u, rec, 1e100, CFL=CFL, sgsIter=sgsIter, maxIter=max_iter,
                                                     tol=tol, fp32Sweeps=fp32)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            if best is None or ms < best[0]:
                best = (ms, n, float(rel[0]))
        states[name] = g.to_host(u)
        out[name] = {"ms": round(best[0], 3), "iterations": best[1], "resRel_rho": best[2],
                     "ms_per_iteration": round(best[0] / best[1], 4)}
    a, b = states["mixed"], states["fp64"]
    out["state_rel_l2_mixed_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["speedup_mixed"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    out["_check"] = oracle_check_data(m, g, ev, u0, types, np.stack([far] * len(types)), False, {}, CFL, sgsIter)
    out["case"] = (f"steady inviscid Mach 0.5 flow over a 0.1H sinusoidal bump, {shape[0]}x{shape[1]}x{shape[2]} "
                   f"= {g.nCell} jittered hexes, p1 VR, Roe, slip wall + free-stream (BCIn) boundaries")
    out["settings"] = {"pseudo_CFL": CFL, "tol_rel_rho": tol, "nRec": 1, "sgsIter": sgsIter, "jacobiCode": 1,
                       "max_iter": max_iter, "physical_dt": "infinite (steady)"}
    return out


def vr_sweep_forms(m, vfv, u, reps=20):
    """Average time of one reconstruction sweep in each form on the benchmark mesh, from the
    library's own per-sweep HIP-event timing: Jacobi and in-place multicolour SOR at p2, and
    the second-order substitute (Gauss rule, least squares) at p1."""
    import torch
    from dndsr_amd import _lib
    from dndsr_amd.cfv import VariationalReconstruction
    L_ = _lib.lib()
    nk = len(_lib.KERNEL_IDS)
    k_vr = _lib.KERNEL_IDS.index("vr_iter")

    def timed(g, fn):
        fn()
        torch.cuda.synchronize()
        _lib.check(L_.dndsr_get_timing(g.ctx.h, None, None, 1))
        _lib.check(L_.dndsr_set_timing(g.ctx.h, 1))
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        _lib.check(L_.dndsr_set_timing(g.ctx.h, 0))
        ms, cnt = np.zeros(nk), np.zeros(nk, np.int64)
        _lib.check(L_.dndsr_get_timing(g.ctx.h, ms.ctypes.data_as(_lib.dp), cnt.ctypes.data_as(_lib.i64p), 1))
        return round(float(ms[k_vr] / max(cnt[k_vr], 1)), 5)

    out = {"mesh": f"{vfv.nCell} periodic hexes (the benchmark mesh)", "timing": "HIP events per sweep, "
           f"average of {reps}"}
    rec, new = vfv.BuildURec(5), vfv.BuildURec(5)
    out["p2_jacobi_ms"] = timed(vfv, lambda: vfv.DoReconstructionIter(rec, new, u))
    out["p2_sor_ms"] = timed(vfv, lambda: vfv.DoReconstructionSOR(rec, u))
    out["p2_sor_colours"] = int(vfv.colors().max()) + 1
    for meth, name in ((1, "gauss"), (2, "least_squares")):
        g1 = VariationalReconstruction(m, dict(maxOrder=1, intOrder=3, subs2ndOrder=meth)).construct()
        r1, n1 = g1.BuildURec(5), g1.BuildURec(5)
        out[f"p1_subs2nd_{name}_ms"] = timed(g1, lambda: g1.DoReconstructionIter(r1, n1, u))
    return out


def laminar_box(n=66, reorder="morton", colour_sort=False):
    """The laminar configuration's mesh and evaluator: the unit box with a 0.1H sin^2 bump on
    the lower wall, the (n+1)^3 jittered lattice points Delaunay-tetrahedralised
    (mesh.bump_box_delaunay, ~6.95 n^3 cells), cells renumbered for locality by the Morton
    order of their centres.  The implicit sweep keeps its own colour-major copy of what it
    reads, so the cells need not be grouped by colour (colour_sort groups them anyway).
    Measured per implicit iteration on the 2.0M-cell mesh (FP64 / FP32 sweeps): Morton
    6.84 / 5.24 ms; Morton grouped by colour 7.97 / 6.45 ms; reverse Cuthill-McKee grouped
    by colour 8.95 / 7.07 ms; Qhull's order grouped by colour 9.33 / 7.61 ms.  Re 1e4 on the
    height and free-stream speed, Mach 0.5, p1 VR (intOrder 3), Roe, laminar viscous fluxes
    (constant mu, Pr 0.72); no-slip adiabatic lower wall, free stream through the Riemann
    flux at inlet, outlet and top, symmetry on the spanwise sides."""
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    gamma, Re = 1.4, 1e4
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / gamma / (gamma - 1) + 0.5 * 0.25])  # Mach 0.5, a = 1, H = 1
    types = [bc.BC_IN, bc.BC_IN, bc.BC_WALL, bc.BC_IN, bc.BC_SYM]
    m = mm.bump_box_delaunay(n)
    if reorder:
        m = mm.reorder_cells(m, reorder)
    vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    if colour_sort:
        m = mm.colour_sorted(m, VariationalReconstruction(m, vs).construct().colors())
    g = VariationalReconstruction(m, vs).construct()
    gas = dict(Rgas=1.0, muGas=0.5 / Re, prGas=0.72, muModel=0)
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * len(types)), viscous=True, idealGasProperty=gas))
    return m, g, ev, far, types, gas


def implicit_laminar_tet(n=66, tol=1e-8, CFL=1000.0, sgsIter=3, max_iter=20000, fixed_iter=0, profile_iter=20):
    """Second half of the metric on the laminar configuration (laminar_box): the implicit
    LU-SGS pseudo-time loop (infinite physical step) to a density residual below tol of its
    running maximum, with FP64 sweeps and with FP32 sweeps (residual, reconstruction and update
    in FP64).  Then profile_iter iterations of each with per-launch kernel timing give the
    rooflines of the loop's kernels.  fixed_iter > 0 runs that many iterations per precision
    instead (short runs for kernel profiles)."""
    import torch
    from dndsr_amd import _lib
    t_mesh = time.perf_counter()
    m, g, ev, far, types, gas = laminar_box(n)
    t_mesh = time.perf_counter() - t_mesh
    u0 = g.to_device(np.tile(far, (g.nCell, 1)))

    def run(fp32, maxIter, tol_, minIter=5):
        u = u0.clone()
        rec = g.BuildURec(5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n_, res, rel, hist = ev.ImplicitEulerStep(u, rec, 1e100, CFL=CFL, sgsIter=sgsIter, maxIter=maxIter,
                                                  minIter=minIter, tol=tol_, fp32Sweeps=fp32)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, n_, float(rel[0]), u

    out = {"case": (f"steady laminar NS, unit box with a 0.1H sin^2 bump, {n + 1}^3 jittered lattice points (10% "
                    f"of the edge, PCG64 seed 20261202) Delaunay-tetrahedralised: {g.nCell} tetrahedra, Re 1e4, "
                    f"Mach 0.5, p1 VR intOrder 3, Roe, no-slip adiabatic wall, BCIn far boundaries, symmetric "
                    f"sides; cells in Morton order ({int(g.colors().max()) + 1} Gauss-Seidel colours)"),
           "cells": g.nCell, "mesh_s": round(t_mesh, 1),
           "settings": {"pseudo_CFL": CFL, "tol_rel_rho": tol, "nRec": 1, "sgsIter": sgsIter, "jacobiCode": 1,
                        "physical_dt": "infinite (steady)"}}
    states = {}
    for name, fp32 in (("fp64", False), ("mixed", True)):
        if fixed_iter:
            ms, n_, rel, u = run(fp32, fixed_iter, 0.0, fixed_iter)
        else:
            run(fp32, 20, 0.0, 20)  # warm-up: kernels, allocations
            print(f"[bench] laminar tet box {name} ...", file=sys.stderr, flush=True)
            ms, n_, rel, u = run(fp32, max_iter, tol)
        states[name] = g.to_host(u)
        out[name] = {"ms": round(ms, 1), "iterations": n_, "resRel_rho": rel, "ms_per_iteration": round(ms / n_, 4),
                     "converged": bool(rel < tol)}
    a, b = states["mixed"], states["fp64"]
    out["state_rel_l2_mixed_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["speedup_mixed"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    if profile_iter and not fixed_iter:
        out["kernels"] = laminar_kernel_rooflines(g, run, profile_iter, sgsIter)
        out["_check"] = oracle_check_data(m, g, ev, u0, types, np.stack([far] * len(types)), True, gas, CFL, sgsIter)
    return out


def oracle_check_data(m, g, ev, u0, types, values, viscous, gas, CFL, sgsIter, iters=2):
    """Two FP64 implicit iterations from the free stream on the device, and what the CPU
    baseline leg needs to repeat them in the oracle (the mesh, boundary table, gas, the
    device's colour order as the serial scan order)."""
    import torch
    from dndsr_amd import bc
    u = u0.clone()
    rec = g.BuildURec(5)
    ev.ImplicitEulerStep(u, rec, 1e100, CFL=CFL, sgsIter=sgsIter, maxIter=iters, minIter=iters, tol=0.0)
    torch.cuda.synchronize()
    col = g.colors()
    return {"mesh": m, "types": types, "values": values, "viscous": viscous, "gas": gas, "CFL": CFL,
            "sgsIter": sgsIter, "iters": iters, "u0": g.to_host(u0), "u": g.to_host(u),
            "zoneDirWeight": bc.euler_dir_weights(types, 1),
            "scanOrder": np.lexsort((np.arange(col.size), col)).astype(np.int64)}


def laminar_kernel_rooflines(g, run, iters, sgsIter):
    """Per-kernel averages of `iters` implicit iterations (per-launch HIP events) with FP64 and
    with FP32 sweeps, against the algorithmic bytes per cell of the tetrahedral p1 case:
    EvaluateRHS (face + viscous + gather passes) B_RHS, the VR sweep B_VR, one SGS sweep over
    all colours 8 (5 nV + 1) + f_c (4 + s (2 + d)) + s (5 nV + 1) with s the sweep's precision
    (rhs, JDiag, the cell's increment in and out; per face the neighbour, its state and
    increment, the face geometry)."""
    from dndsr_amd import _lib
    L_ = _lib.lib()
    nk = len(_lib.KERNEL_IDS)
    N = g.nCell
    f_c = g.nFace / N
    b_vr, b_rhs = alg_bytes(3, 3, 5, 4, f_c, g.nGFace)
    res = {"bytes_per_cell": {"eval_rhs": round(b_rhs, 1), "vr_iter": b_vr}}
    for name, fp32 in (("fp64", False), ("mixed", True)):
        _lib.check(L_.dndsr_get_timing(g.ctx.h, None, None, 1))
        _lib.check(L_.dndsr_set_timing(g.ctx.h, 1))
        run(fp32, iters, 0.0, iters)
        _lib.check(L_.dndsr_set_timing(g.ctx.h, 0))
        ms, cnt = np.zeros(nk), np.zeros(nk, np.int64)
        _lib.check(L_.dndsr_get_timing(g.ctx.h, ms.ctypes.data_as(_lib.dp), cnt.ctypes.data_as(_lib.i64p), 1))
        k = {nm: (float(ms[i]), int(cnt[i])) for i, nm in enumerate(_lib.KERNEL_IDS) if cnt[i]}
        per_it = lambda nm: k.get(nm, (0.0, 0))[0] / iters  # noqa: E731
        rhs_ms = per_it("rhs_face") + per_it("rhs_viscous") + per_it("rhs_gather")
        vr_ms = per_it("vr_iter")
        # a symmetric sweep after the first skips its first colour (the previous sweep's
        # last, whose increments it would repeat): time per full pass over the mesh
        sweeps = 2 * (1 + sgsIter)
        cs = np.bincount(np.asarray(g.colors()))
        processed = sweeps * N - (1 + sgsIter) * int(cs[-1]) - sgsIter * int(cs[0])
        sgs_ms = per_it("sgs_colour") * N / processed
        s_ = 4 if fp32 else 8
        b_sgs = 8 * (5 * 5 + 1) + f_c * (4 + s_ * (2 + 3)) + s_ * (5 * 5 + 1)
        ent = {}
        for key, t, b in (("eval_rhs", rhs_ms, b_rhs), ("vr_iter", vr_ms, b_vr), ("sgs_sweep", sgs_ms, b_sgs)):
            if t > 0:
                gbs = N * b / (t * 1e-3) / 1e9
                ent[key] = {"avg_ms": round(t, 4), "bytes_per_cell": round(b, 1), "achieved": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 4)}
        ent["passes_ms_per_iteration"] = {nm: round(v[0] / iters, 4) for nm, v in k.items()}
        ent["sgs_cells_per_iteration"] = {"processed": processed, "full_sweeps": sweeps * N}
        res[name] = ent
    return res


def loop_ms(fn, reps=20):
    """Average milliseconds of one call over `reps` back-to-back calls bracketed by one event
    pair on torch's current stream, the stream the library launches on (no per-launch events)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def stream_peaks(ctx, gib=1.0, reps=10):
    """The device's own read stream (every element read once) and copy over gib GiB
    (dndsr_stream_probe), in GB/s."""
    import ctypes
    import torch
    from dndsr_amd import _lib
    n = int(gib * (1 << 30) // 8) // 2 * 2
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    part = torch.empty(4096, dtype=torch.float64, device="cuda")
    L_ = _lib.lib()
    ms = ctypes.c_double()
    _lib.check(L_.dndsr_stream_probe(ctx.h, 0, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(part.data_ptr()), n,
                                     reps, ctypes.byref(ms)))
    read = n * 8 / (ms.value * 1e-3) / 1e9
    _lib.check(L_.dndsr_stream_probe(ctx.h, 1, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), n,
                                     reps, ctypes.byref(ms)))
    copy = 2 * n * 8 / (ms.value * 1e-3) / 1e9
    del x, y, part
    return {"read_GBs": round(read, 1), "copy_GBs": round(copy, 1), "bytes": n * 8, "passes": reps}


def _converge(ev, u, uRec, absTol, chunk=1000, maxChunks=60, log=None, **settings):
    """Steady iterations (infinite physical step) in chunks until the density residual is at
    most absTol, or has stopped falling (rounding level) within 100 absTol; returns
    (iterations, last residual norms, seconds)."""
    import torch
    t0, nTot, res, hist = time.perf_counter(), 0, None, []
    for _ in range(maxChunks):
        n, res, _, _ = ev.ImplicitEulerStep(u, uRec, 1e100, maxIter=chunk, minIter=chunk, tol=0.0, **settings)
        nTot += n
        hist.append(float(res[0]))
        if log:
            log(f"  {nTot} iterations, density residual {res[0]:.3e}")
        if res[0] <= absTol or (len(hist) > 2 and hist[-1] > 0.7 * hist[-3] and hist[-1] < 100 * absTol):
            break
    torch.cuda.synchronize()
    return nTot, res, time.perf_counter() - t0


def adjoint_bump(nx=516, ny=258, Re=5e3, M=0.7, fd=True, h=1e-6, seed=20261203):
    """The 2-D discrete-adjoint workload: Mach 0.7, Re 5e3 laminar flow over a lower-wall bump
    of eight Hicks-Henne functions (coefficients uniform in [-0.01, 0.01], seeded) in the
    channel [-1, 2] x [0, 1.5] of 1.5 nx ny mixed triangles and quadrilaterals (199692 cells
    at 516 x 258); drag = x-force on the wall.  The steady state is converged with the
    block Roe-Jacobian LU-SGS (CFL 1000, 5 extra symmetric sweeps) to a density residual of
    its rounding level (about 6e-13 of the free-stream start); the adjoint is solved by
    GMRES(300) on the transposed system, preconditioned by transposed block LU-SGS sweeps, to
    1e-11; the shape gradient comes from the Lagrangian on moved meshes (step h).  With fd,
    each of the 8 gradient entries is checked against central differences (step h) of the
    drag of steady states converged on the moved meshes to the same absolute residual,
    started from the unmoved steady state."""
    import torch
    from dndsr_amd import bc
    from dndsr_amd.adjoint import SteadyAdjoint, zone_force, zone_force_vjp
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    from dndsr_amd.shape import HicksHenneBumpChannel, shape_gradient

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ch = HicksHenneBumpChannel(nx, ny)
    alpha = np.random.default_rng(seed).uniform(-0.01, 0.01, 8)
    vals, gas = ch.flow(M=M, Re=Re)
    types = [bc.bc_code(t) for t in ch.BC_TYPES]
    solver = dict(CFL=1000.0, sgsIter=5)
    adjSolver = dict(CFL=1000.0, sgsIter=5, restart=300, maxIter=6000, tol=1e-11)

    def build(a, roeJ=False):
        g = VariationalReconstruction(ch.mesh(a), dict(maxOrder=1, intOrder=3, bcTypes=types)).construct()
        return EulerEvaluator(g, dict(gamma=1.4, riemannSolverType="Roe", bcTypes=types, bcValues=vals,
                                      viscous=True, idealGasProperty=gas, useRoeJacobian=roeJ))

    def drag(ev_, u, rec, dlc=None):
        return zone_force(ev_, u, rec, 0, dlc)[0]

    out = {"workload": f"Hicks-Henne bump channel, {ch.base.nCell} mixed tri/quad cells, Mach {M}, Re {Re:g}, "
                       "p1 VR, Roe, laminar", "nCell": ch.base.nCell, "alpha": [round(float(a), 6) for a in alpha]}
    ev = build(alpha, roeJ=True)
    g = ev.vfv
    u = g.to_device(np.tile(vals[1], (g.nCell, 1)))
    rec = g.BuildURec(5)
    n0, res0, _, _ = ev.ImplicitEulerStep(u, rec, 1e100, maxIter=1, minIter=1, tol=0.0, **solver)
    absTol = 1e-13 * res0[0]  # below the rounding level: the solves stop where the residual stops falling
    log(f"[adjoint bump] forward solve, target density residual {absTol:.3e}")
    n, res, t = _converge(ev, u, rec, absTol, log=log, **solver)
    dTau = g.BuildUDof(1)[0]
    ev.EvaluateDt(dTau, u, None, CFL=1.0)
    J0 = drag(ev, u, rec)
    out["forward"] = {"iterations": n + n0, "seconds": round(t, 2), "density_residual": float(res[0]),
                      "density_residual_start": float(res0[0]), "drag": J0}
    adj = SteadyAdjoint(ev, CFL=adjSolver["CFL"], sgsIter=adjSolver["sgsIter"], block=True)
    uB, rB = zone_force_vjp(ev, u, rec, 0, [1.0, 0.0, 0.0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam, mu, nIt = adj.solve(u, rec, uB, rB, tol=adjSolver["tol"], maxIter=adjSolver["maxIter"], method="gmres",
                             restart=adjSolver["restart"])
    torch.cuda.synchronize()
    tAdj = time.perf_counter() - t0
    log(f"[adjoint bump] adjoint: {nIt} GMRES iterations, {tAdj:.2f} s")
    t0 = time.perf_counter()
    gS = shape_gradient(adj, u, rec, lam, mu, drag, build, alpha, h=h)
    tShape = time.perf_counter() - t0
    out["adjoint"] = {"method": "GMRES(300), preconditioner 6 symmetric transposed Roe-Jacobian block LU-SGS sweeps "
                                "(CFL 1000)", "iterations": nIt,
                      "seconds": round(tAdj, 2), "residual_rel": adj.history[-1] / max(adj.history[0], 1e-300),
                      "shape_gradient_seconds": round(tShape, 2), "gradient": [float(x) for x in gS]}
    if fd:
        gF = np.zeros(8)
        tF = time.perf_counter()
        for k in range(8):
            Js = []
            for sgn in (1.0, -1.0):
                a = alpha.copy()
                a[k] += sgn * h
                evk = build(a, roeJ=True)
                uk, rk = u.clone(), rec.clone()
                nk, resk, tk = _converge(evk, uk, rk, absTol, **solver)
                evk.EvaluateDt(dTau, uk, None, CFL=1.0)
                Js.append(drag(evk, uk, rk))
                log(f"[adjoint bump] fd {k} {sgn:+.0f}: {nk} iterations, residual {resk[0]:.3e}, {tk:.1f} s")
                evk.ctx.close()
            gF[k] = (Js[0] - Js[1]) / (2 * h)
        err = np.abs(gS - gF) / np.abs(gF).max()
        out["fd"] = {"step": h, "gradient": [float(x) for x in gF], "max_rel_error": float(err.max()),
                     "rel_error": [float(x) for x in err], "seconds": round(time.perf_counter() - tF, 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-implicit", action="store_true", help="skip the implicit time-to-tolerance measurements")
    ap.add_argument("--tet-iterations", type=int, default=0, metavar="K",
                    help="only run K implicit iterations of the laminar tetrahedral case with FP64 and with FP32 "
                         "sweeps (a short run for kernel profiles)")
    ap.add_argument("--adjoint-bump", action="store_true",
                    help="only run the 2-D adjoint workload (Hicks-Henne bump drag gradient); with --adjoint-fd also "
                         "the central-difference check of its 8 entries")
    ap.add_argument("--adjoint-fd", action="store_true")
    ap.add_argument("--adjoint-step", type=float, default=1e-6, help="parameter step of the shape derivatives")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "launches": 3}))
        return
    if args.adjoint_bump:
        print(json.dumps(adjoint_bump(fd=args.adjoint_fd, h=args.adjoint_step)), flush=True)
        return
    if args.tet_iterations:
        r = implicit_laminar_tet(fixed_iter=args.tet_iterations)
        print(json.dumps(r), flush=True)
        return

    import torch
    from dndsr_amd import _lib, fields
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator

    n = args.n
    L = 2 * np.pi
    t_setup = time.perf_counter()
    m = mm.periodic_hex_grid(n, n, n, L, L, L)
    vfv = VariationalReconstruction(m, dict(maxOrder=2, intOrder=5)).construct()
    ev = EulerEvaluator(vfv, dict(gamma=1.4, riemannSolverType="HLLEP"))
    u0 = vfv.cell_average(lambda p: fields.taylor_green(p, 0.1, 1.4), 5)
    u = vfv.to_device(u0)
    uRec = vfv.BuildURec(5)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    N = vfv.nCell
    CFL = 0.5

    # the frhs of the initial state, for the check against the oracle
    rec1, scratch = vfv.BuildURec(5), vfv.BuildURec(5)
    vfv.DoReconstructionIter(scratch, rec1, u)
    dt0 = vfv.BuildUDof(1)[0]
    ev.EvaluateDt(dt0, u, None, CFL=CFL, UseLocaldt=False)
    rhs0 = vfv.BuildUDof(5)
    ev.EvaluateRHS(rhs0, None, u, rec1, rec1)
    checks = {"tgv": {"u0": u0, "rhs": vfv.to_host(rhs0)}}

    ev.SSPRK3(u, uRec, args.warmup, CFL, 1)
    torch.cuda.synchronize()
    L_ = _lib.lib()
    _lib.check(L_.dndsr_get_timing(vfv.ctx.h, None, None, 1))
    _lib.check(L_.dndsr_set_timing(vfv.ctx.h, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev.SSPRK3(u, uRec, args.steps, CFL, 1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.check(L_.dndsr_set_timing(vfv.ctx.h, 0))
    nk = len(_lib.KERNEL_IDS)  # == DNDSR_NUM_KERNEL_IDS: dndsr_get_timing fills that many entries
    ms = np.zeros(nk)
    cnt = np.zeros(nk, np.int64)
    _lib.check(L_.dndsr_get_timing(vfv.ctx.h, ms.ctypes.data_as(_lib.dp), cnt.ctypes.data_as(_lib.i64p), 1))
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / max(args.steps, 1)
    value = N * 3 * args.steps / elapsed

    un = vfv.to_host(u)
    if not np.all(np.isfinite(un)):
        raise RuntimeError("non-finite state after the timed steps")

    b_vr, b_rhs = alg_bytes(3, vfv.M, 5, 6, 3, vfv.nGFace)
    kern = {}
    names = _lib.KERNEL_IDS
    for k, nm in enumerate(names):
        if cnt[k]:
            kern[nm] = {"ms_total": float(ms[k]), "launches": int(cnt[k]), "avg_ms": float(ms[k] / cnt[k])}
    vr_avg = kern["vr_iter"]["avg_ms"]
    rhs_avg = kern["rhs_face"]["avg_ms"] + kern["rhs_gather"]["avg_ms"]

    # the same kernels as back-to-back loops under one event pair (no per-launch events)
    rr, rn = vfv.BuildURec(5), vfv.BuildURec(5)
    vfv.DoReconstructionIter(rn, rr, u)
    rhs_l = vfv.BuildUDof(5)
    loop_rhs = loop_ms(lambda: ev.EvaluateRHS(rhs_l, None, u, rr, rr, check_finite=False))
    loop_vr = loop_ms(lambda: vfv.DoReconstructionIter(rr, rn, u))
    peaks = stream_peaks(vfv.ctx)

    def entry(b, t, t_loop, tag):
        a = N * b / (t * 1e-3) / 1e9
        a_l = N * b / (t_loop * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_cell": b,
                "avg_ms": t, "frac": a / HBM_PEAK_GBS, "traffic": read_traffic(tag),
                "loop_avg_ms": t_loop, "loop_achieved": a_l, "loop_frac": a_l / HBM_PEAK_GBS,
                "frac_of_measured_read": a / peaks["read_GBs"]}

    roof = {"vr_iter": entry(b_vr, vr_avg, loop_vr, "vr_iter_tgv64_p2"),
            "eval_rhs": entry(b_rhs, rhs_avg, loop_rhs, "rhs_tgv64_p2")}
    dom = "vr_iter" if kern["vr_iter"]["ms_total"] >= kern["rhs_face"]["ms_total"] + kern["rhs_gather"]["ms_total"] \
        else "eval_rhs"
    droof = roof[dom]
    roofline = {"bound": "hbm", "achieved": round(droof["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(droof["frac"], 4), "traffic": droof["traffic"], "kernel": dom,
                "bytes_per_launch": N * droof["bytes_per_cell"], "avg_launch_ms": round(droof["avg_ms"], 5),
                "measured_peaks": peaks}

    forms = vr_sweep_forms(m, vfv, u)

    implicit = implicit_tet = None
    if not args.no_implicit:
        implicit = implicit_bump_time_to_tol()
        checks["bump"] = implicit.pop("_check")
        implicit_tet = implicit_laminar_tet()
        if "_check" in implicit_tet:
            checks["laminar"] = implicit_tet.pop("_check")

    cpu = errors = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(checks=checks)
        errors = cpu.pop("error_vs_oracle")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3-D Euler Taylor-Green vortex M0.1, periodic [0,2pi]^3, {n}^3 hex = {N} cells, "
                               f"p2 VR (9 modes), intOrder 5, HLLE+, FP64, SSP-RK3 CFL {CFL}, "
                               f"1 Jacobi VR sweep per RHS", "cells": N, "rhs_evals_per_step": 3},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "error_vs_oracle": errors,
        "implicit": implicit,
        "implicit_laminar_tet": implicit_tet,
        "vr_sweep_forms": forms,
        "kernels": kern,
        "kernel_timing": "kernels: per-launch HIP event pairs during the timed steps; loop_*: back-to-back calls "
                         "under one event pair",
        "roofline_all": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in r.items()}
                         for k, r in roof.items()},
        "setup_s": round(t_setup, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
