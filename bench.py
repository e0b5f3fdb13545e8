"""Benchmark of the compact-FV hot path on one MI355X.

Workload (the reference-path configuration the metric is quoted on): 3-D Euler
Taylor-Green vortex, Mach 0.1, periodic [0, 2 pi]^3 with 64^3 = 262144 uniform
hexahedra, 3rd-order VR (p2, 9 modes per variable, intOrder 5), HLLE+ flux, FP64,
explicit SSP-RK3 at CFL 0.5.  One step = EvaluateDt + 3 x (one Jacobi VR sweep +
EvaluateRHS + stage combination), all on the device (dndsr_ssprk3).

value = cell-updates per second per RHS evaluation = nCell * 3 * steps / time.
The roofline entry is the dominant kernel: algorithmic bytes per launch divided
by its average launch time, measured with HIP events on its stream during the
timed region.  cpu_baseline times the CPU oracle on the same 64^3 workload with all
physical host cores and with one core (median of five repetitions).
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


def cpu_baseline(n=64, reps=5, evals_all=4, evals_one=1):
    """The reference-path CPU baseline: the oracle's frhs (one Jacobi VR sweep + EvaluateRHS, the
    OpenMP face-buffer + cell-gather form of EvaluateRHS.hxx:153-159,640-662) on the benchmark
    configuration itself (TGV n^3 hex, p2, HLLE+, FP64), timed with all physical cores available
    and with one core: after one warm-up evaluation, the median of `reps` repetitions of
    `evals_all` (all cores) or `evals_one` (one core) evaluations; value = cells / median seconds
    per evaluation.  The oracle's setup (metrics and VR coefficients) is not timed."""
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
    return {"value": v_all, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "value_1core": v_one, "cores_used_1core": 1,
            "sample": (f"the benchmark workload itself (TGV {n}^3 = {o.nCell} hex, p2, HLLE+, FP64): oracle frhs = "
                       f"Jacobi VR sweep + EvaluateRHS (OpenMP face buffer + cell gather); median of {reps} "
                       f"repetitions of {evals_all} evaluations on {cores} cores "
                       f"({res['all'][0] * 1e3:.1f} ms each) and of {evals_one} on 1 core "
                       f"({res['one'][0] * 1e3:.1f} ms); oracle setup {t_setup:.1f} s untimed"),
            "host": info}


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


def implicit_laminar_tet(n=69, Re=1e4, tol=1e-8, CFL=1000.0, sgsIter=3, max_iter=20000, fixed_iter=0):
    """Second half of the metric on the laminar configuration: steady laminar
    Navier-Stokes flow in the unit box with a sinusoidal bump of 0.1 of the height on
    the lower wall (y_w = 0.1 sin^2(pi x)), an n^3 lattice (n = 69: 70^3 vertices,
    jittered uniformly by +-10 % of the edge, PCG64 seed 20261202) split into six
    tetrahedra per cube (1 971 054 cells), Re = 1e4 on the height and the free-stream
    speed, Mach 0.5, p1 VR (intOrder 3: 6-point face and 8-point cell rules), Roe,
    laminar viscous fluxes (constant mu, Pr 0.72), no-slip adiabatic lower wall, free
    stream through the Riemann flux at inlet, outlet and top, spanwise periodic; the
    implicit LU-SGS pseudo-time loop (infinite physical step) to a density residual
    below tol of its running maximum, with FP64 sweeps and with FP32 sweeps (residual,
    reconstruction and update in FP64).  Cells are grouped by Gauss-Seidel colour
    (mesh.colour_sorted) first.  fixed_iter > 0 runs that many iterations per precision
    instead (short runs for kernel profiles)."""
    import torch
    from dndsr_amd import bc
    from dndsr_amd import mesh as mm
    from dndsr_amd.cfv import VariationalReconstruction
    from dndsr_amd.euler import EulerEvaluator
    gamma = 1.4
    far = np.array([1.0, 0.5, 0.0, 0.0, 1 / gamma / (gamma - 1) + 0.5 * 0.25])  # Mach 0.5, a = 1, H = 1
    types = [bc.BC_IN, bc.BC_IN, bc.BC_WALL, bc.BC_IN]
    m = mm.bump_box_tet(n)
    vs = dict(maxOrder=1, intOrder=3, bcTypes=types)
    m = mm.colour_sorted(m, VariationalReconstruction(m, vs).construct().colors())  # coalesced colour passes
    g = VariationalReconstruction(m, vs).construct()
    ev = EulerEvaluator(g, dict(gamma=gamma, riemannSolverType="Roe", bcTypes=types,
                                bcValues=np.stack([far] * len(types)), viscous=True,
                                idealGasProperty=dict(Rgas=1.0, muGas=0.5 / Re, prGas=0.72, muModel=0)))
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

    out = {"case": (f"steady laminar NS, unit box with a 0.1H sin^2 bump, {n}^3 cubes x 6 = {g.nCell} jittered "
                    f"tetrahedra, Re {Re:g}, Mach 0.5, p1 VR intOrder 3, Roe, no-slip adiabatic wall, BCIn far "
                    f"boundaries, cells grouped by colour"),
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
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
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
    if args.tet_iterations:
        r = implicit_laminar_tet(fixed_iter=args.tet_iterations)
        print(json.dumps(r), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

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
    roof = {
        "vr_iter": {"bound": "hbm", "achieved": N * b_vr / (vr_avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "bytes_per_cell": b_vr, "avg_ms": vr_avg,
                    "traffic": read_traffic("vr_iter_tgv64_p2")},
        "eval_rhs": {"bound": "hbm", "achieved": N * b_rhs / (rhs_avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "bytes_per_cell": b_rhs, "avg_ms": rhs_avg,
                     "traffic": read_traffic("rhs_tgv64_p2")},
    }
    for r in roof.values():
        r["frac"] = r["achieved"] / r["peak"]
    dom = "vr_iter" if kern["vr_iter"]["ms_total"] >= kern["rhs_face"]["ms_total"] + kern["rhs_gather"]["ms_total"] \
        else "eval_rhs"
    droof = roof[dom]
    roofline = {"bound": "hbm", "achieved": round(droof["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(droof["frac"], 4), "traffic": droof["traffic"], "kernel": dom,
                "bytes_per_launch": N * droof["bytes_per_cell"], "avg_launch_ms": round(droof["avg_ms"], 5)}

    forms = vr_sweep_forms(m, vfv, u)

    implicit = implicit_tet = None
    if not args.no_implicit:
        implicit = implicit_bump_time_to_tol()
        implicit_tet = implicit_laminar_tet()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline()

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3-D Euler Taylor-Green vortex M0.1, periodic [0,2pi]^3, {n}^3 hex = {N} cells, "
                               f"p2 VR (9 modes), intOrder 5, HLLE+, FP64, SSP-RK3 CFL {CFL}, "
                               f"1 Jacobi VR sweep per RHS", "cells": N, "rhs_evals_per_step": 3},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "implicit": implicit,
        "implicit_laminar_tet": implicit_tet,
        "vr_sweep_forms": forms,
        "kernels": kern,
        "roofline_all": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in r.items()}
                         for k, r in roof.items()},
        "setup_s": round(t_setup, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
