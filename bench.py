"""Benchmark: AIC assembly + solve of a 30,000-panel lifting body on one MI355X.

Workload: blended-wing-body loft, 12 random NACA 4-digit sections (seed 5),
250 chordwise x 120 spanwise vortex-ring quads (30,000 panels, 120 trailing-edge
shedding edges, rigid wake along the freestream), AoA 5 deg.  One step is one
`solve(body, Uinfs, Das, Dbs)` from device-resident geometry and freestream to
FP64 strengths on the device:

  mixed  = AIC computed in FP64 and stored in FP32 (plus the FP64 copy for the
           residual); restarted GMRES on the FP32 AIC, preconditioned by block-Jacobi
           (250-panel blocks) and a Galerkin coarse correction on chordwise-strip
           indicators, inside FP64 iterative refinement                  -> `value`
  fp64   = FP64 AIC assembly + FP64 LU (getrf/getrs)                   -> reported beside it

The Cp error of the mixed result is measured against the FP64 LU result at the
control points (the FP64 path itself is checked against the CPU oracle by the
parity tests at smaller sizes).  One JSON line is printed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per k_gemv<float> launch at 30,000 panels from rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by the
# same counters on a 1 GiB plain copy (FETCH x2.0, WRITE x1.0): profiles/gemv_f32_bwb30k_hbm_counters.md
GEMV_TRAFFIC_30K = 3.6052e9
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_body(fp, nc, ns, alpha):
    grid, sh = fp.geometry.bwb(seed=5, n_chord=nc, n_span=ns)
    body = fp.RigidWakeBody(grid, sh)
    V = fp.direction(alpha)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    return body, U, D


def cpu_baseline(fp, body, U, D, seconds_hint=20.0):
    """Restated reference CPU path (the oracle) on a bounded sample.

    Assembly: oracle `_G_Uvortexring!` restatement (OpenMP over panel columns)
    for the first R control points of the same body, scaled to all rows.
    Solve: LAPACK getrf+getrs (scipy/OpenBLAS, all host threads) on an n x n
    slice, scaled by (N/n)^3.
    """
    import types
    import scipy.linalg as sla
    from oracle import oracle as orc
    b = types.SimpleNamespace(nodes=body.grid.nodes, cell_ptr=body.grid.cell_ptr, cell_idx=body.grid.cell_idx,
                              ncells=body.ncells, element="vortexring", shedding=body.shedding,
                              nsheddings=body.nsheddings, kerneloffset=body.kerneloffset,
                              kernelcutoff=body.kernelcutoff, CPoffset=body.CPoffset, charlen_kind=0)
    N = body.ncells
    normals, cps, _ = orc.geometry(b.nodes, b.cell_ptr, b.cell_idx, b.CPoffset, 0)
    R = 1200
    t = time.perf_counter()
    Gs, _, _ = orc.assemble([b], [D], [D], targets=cps[:, :R], normals=normals[:, :R])
    t_asm = (time.perf_counter() - t) * N / R
    n = 5000
    A = np.asfortranarray(np.vstack([Gs[:, :n]] * ((n + R - 1) // R))[:n, :n] + np.eye(n))
    t = time.perf_counter()
    lu = sla.lu_factor(A, overwrite_a=True, check_finite=False)
    sla.lu_solve(lu, np.ones(n), check_finite=False)
    t_lu = (time.perf_counter() - t) * (N / n) ** 3
    cores = os.cpu_count()
    return {"value": (t_asm + t_lu) * 1e3, "unit": "ms", "cores": cores, "kind": "port",
            "sample": "oracle AIC of the first %d of %d control points (scaled x%.1f) + LAPACK getrf/getrs "
                      "(OpenBLAS) on a %dx%d slice (scaled (N/n)^3); %d host threads"
                      % (R, N, N / R, n, n, cores),
            "assembly_ms_est": t_asm * 1e3, "lu_ms_est": t_lu * 1e3}


def vendor_lu(fp, body, N):
    """rocSOLVER FP64 getrf+getrs (through torch) on the same FP64 AIC: an external anchor for the FP64 path."""
    try:
        import torch
        if not torch.cuda.is_available():
            return {"error": "torch sees no GPU (device_count %d)" % torch.cuda.device_count()}
        G = fp.assemble(body, "f64")
        A = torch.from_numpy(G).cuda()
        del G
        b = torch.ones(N, 1, dtype=torch.float64, device="cuda")
        torch.linalg.lu_factor(A[:2048, :2048].contiguous())
        torch.cuda.synchronize()
        t = time.perf_counter()
        LU, piv = torch.linalg.lu_factor(A)
        x = torch.linalg.lu_solve(LU, piv, b)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        del A, LU, piv, x
        torch.cuda.empty_cache()
        return {"ms": round(ms, 2), "what": "torch.linalg.lu_factor+lu_solve (rocSOLVER) FP64, matrix resident"}
    except Exception as e:  # the anchor is optional
        return {"error": str(e)[:200]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chord", type=int, default=250)
    ap.add_argument("--span", type=int, default=120)
    ap.add_argument("--block", type=int, default=250, help="block-Jacobi block size")
    ap.add_argument("--coarse", type=int, default=-1, help="coarse chunk (default: one chordwise strip)")
    ap.add_argument("--tol", type=float, default=1e-10, help="FP64 relative residual target of the mixed solve")
    ap.add_argument("--inner-tol", type=float, default=1e-5, help="relative tolerance of each inner FP32 GMRES")
    ap.add_argument("--vendor", action="store_true", help="also time rocSOLVER FP64 getrf through torch")
    ap.add_argument("--mixed-only", action="store_true", help="profiling runs: mixed steps only (no FP64 LU, no CPU)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="profiling runs: also run a plain 1 GiB device copy (torch) to calibrate HBM counters")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    fp = __graft_entry__.load_package()
    ctx = fp.context()
    alpha = 5.0
    log("bench: building BWB %d x %d" % (args.chord, args.span))
    body, U, D = build_body(fp, args.chord, args.span, alpha)
    N = body.ncells
    fp.solvers.set_freestream(body, U, D, D)
    coarse = args.chord if args.coarse < 0 else args.coarse
    mixed_opts = {"precond_block": args.block, "coarse_block": coarse, "tol": args.tol, "inner_tol": args.inner_tol,
                  "restart": 100, "residual_mode": 1}

    def run(solver, opts):
        return fp.solve(body, U, D, D, solver=solver, solver_optargs=opts)

    for _ in range(args.warmup):
        log("bench: warmup mixed")
        run(fp.solve_gmres_ir, mixed_opts)
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    stats = []
    for k in range(args.steps):
        stats.append(run(fp.solve_gmres_ir, mixed_opts))
        log("bench: mixed step %d: %s" % (k, stats[-1]))
    ctx.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps * 1e3
    prof = {k: ctx.profile_get(k) for k in ("gemv_f32", "gemv_f64", "assemble_vr_f64", "assemble_wake", "bj_factor",
                                           "bj_invert", "bj_apply", "coarse_setup", "coarse_apply")}
    ctx.profile(False)
    gam_mixed = body.strength.copy()
    Us_m = fp.calcfield_U(body, body)
    cp_m = fp.calcfield_Cp(body, 1.0)
    if args.copy_calib:
        ctx.check(fp._native.lib().fp_ctx_copy_calibrate(ctx.handle, 1 << 30))
        log("bench: copy calibration: 1 GiB read + 1 GiB written (__amd_rocclr_copyBuffer)")
    if args.mixed_only:
        launches, ms = prof["gemv_f32"]
        log("bench: gemv_f32 %d launches avg %.4f ms" % (launches, ms / max(launches, 1)))
        print(json.dumps({"mixed_ms": round(t_mixed, 3), "stats": stats[-1]}), flush=True)
        return

    log("bench: fp64 LU")
    ctx.profile(True)
    t64 = []
    st64 = None
    for k in range(args.fp64_steps + 1):
        ctx.synchronize()
        t = time.perf_counter()
        st64 = run(fp.solve_ludiv, {})
        ctx.synchronize()
        t64.append((time.perf_counter() - t) * 1e3)
        log("bench: fp64 step %d: %.1f ms %s" % (k, t64[-1], st64))
    t_fp64 = min(t64[1:]) if len(t64) > 1 else t64[0]
    prof64 = {k: ctx.profile_get(k) for k in ("assemble_vr_f64", "gemm_f64", "panel_tslu", "trsm", "getrs")}
    ctx.profile(False)
    gam64 = body.strength.copy()
    fp.calcfield_U(body, body)
    cp64 = fp.calcfield_Cp(body, 1.0)
    cp_err = float(np.abs(cp_m - cp64).max())
    gam_err = float(np.abs(gam_mixed - gam64).max() / np.abs(gam64).max())

    # roofline of the dominant kernel of the mixed step (FP32 GEMV of the Krylov matvec)
    launches, ms = prof["gemv_f32"]
    avg_ms = ms / max(launches, 1)
    gemv_bytes = 4.0 * N * N + 16.0 * N
    achieved = gemv_bytes / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    roof = {"bound": "hbm", "kernel": "k_gemv<float>", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": GEMV_TRAFFIC_30K if N == 30000 else None,
            "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "algorithmic_bytes_per_launch": gemv_bytes}
    kernels = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in prof.items()}
    kernels_fp64 = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in prof64.items()}

    vendor = None
    if args.vendor:
        vendor = vendor_lu(fp, body, N)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(fp, body, U, D)

    line = {
        "metric": "AIC+solve time at 30k panels, mixed+refine vs FP64 @ Cp err<=1e-6; adjoint ms",
        "value": round(t_mixed, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mixed, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32 AIC + f64 refinement", "data": "synthetic",
        "config": {"workload": "bwb-%dx%d vortex-ring + rigid wake, AoA %.1f, GMRES-IR (FP32 AIC, block-Jacobi %d)"
                               % (args.chord, args.span, alpha, args.block),
                   "panels": N, "shedding_edges": body.nsheddings},
        "fp64_lu_ms": round(t_fp64, 3), "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
        "cp_err_max": cp_err, "gamma_rel_err": gam_err,
        "mixed_stats": stats[-1], "fp64_stats": st64, "adjoint_ms": None,
        "kernels": kernels, "kernels_fp64": kernels_fp64, "vendor_fp64_getrf": vendor, "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
