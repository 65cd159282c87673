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
control points (`cp_err_max`) and against the oracle's FP64 LAPACK solution of
the same body (`vs_oracle`, the committed fixture tests/golden/bwb30k_oracle.npz).

Adjoint leg: dCL/dnodes, dCDi/dnodes and dCL/dUinf, dCDi/dUinf of a
10,000-panel wing (NACA 2412, AR 8, linear twist 0 -> -3 deg, 100 x 100 quads,
AoA 4 deg) by the offset-probe lift/drag objective: FP32 forward solve with FP64
refinement, then the hand-written reverse pass (transpose solve in the same
mixed solver).  Its gradient is checked against central finite differences
(h = 1e-6) of the FP64-LU forward on 64 seeded node coordinates and all three
Uinf components.  One JSON line is printed.
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
# same counters on a 1 GiB plain copy (FETCH x2.0, WRITE x1.0)
GEMV_TRAFFIC_30K = 3.6040e9
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


ORACLE_FIXTURE = os.path.join(ROOT, "tests", "golden", "bwb30k_oracle.npz")


def oracle_fixture_errors(fp, body, gamma):
    """Error of the mixed result against the oracle's FP64 LAPACK solution of the same 30,000-panel body
    (strengths and Cp without Kutta averaging at the control points, stored as data by
    tests/golden/make_bwb_fixture.py); None when the body is not the fixture's."""
    import hashlib
    if not os.path.exists(ORACLE_FIXTURE):
        return None
    ref = np.load(ORACLE_FIXTURE)
    h = hashlib.sha256(np.ascontiguousarray(body.grid.nodes, dtype=np.float64).tobytes()).hexdigest()
    if h != str(ref["node_sha256"]):
        return None
    cp = fp.calcfield_Cp(body, 1.0, correct_kuttacondition=False)
    return {"cp_err_max": float(np.abs(cp - ref["cp"]).max()),
            "gamma_rel_err": float(np.abs(gamma - ref["gamma"]).max() / np.abs(ref["gamma"]).max()),
            "fixture": "tests/golden/bwb30k_oracle.npz (oracle AIC + LAPACK getrf/getrs, FP64)"}


def cpu_baseline(fp, body, U, D, full=False):
    """Restated reference CPU path (the oracle) on the host cores.

    Assembly: the oracle's `_G_Uvortexring!` restatement (OpenMP over panel
    columns) of the whole 30,000 x 30,000 FP64 AIC with its wake columns.
    Solve: LAPACK getrf + getrs (scipy/OpenBLAS, all host threads) of that AIC
    (``full``, the default: about 10-15 s on 16 threads), or of its 10,000 x
    10,000 leading block scaled by (N/n)^3.
    """
    import types
    import scipy.linalg as sla
    from oracle import oracle as orc
    b = types.SimpleNamespace(nodes=body.grid.nodes, cell_ptr=body.grid.cell_ptr, cell_idx=body.grid.cell_idx,
                              ncells=body.ncells, element="vortexring", shedding=body.shedding,
                              nsheddings=body.nsheddings, kerneloffset=body.kerneloffset,
                              kernelcutoff=body.kernelcutoff, CPoffset=body.CPoffset, charlen_kind=0)
    N = body.ncells
    t = time.perf_counter()
    G, normals, _ = orc.assemble([b], [D], [D])
    t_asm = time.perf_counter() - t
    n = N if full else min(N, 10000)
    A = np.asfortranarray(G[:n, :n]) if n < N else G
    del G
    rhs = -np.einsum("ij,ij->j", U[:, :n], normals[:, :n])
    t = time.perf_counter()
    lu = sla.lu_factor(A, overwrite_a=True, check_finite=False)
    sla.lu_solve(lu, rhs, check_finite=False)
    t_lu = (time.perf_counter() - t) * (N / n) ** 3
    del A, lu
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    sample = ("oracle AIC of all %d control points (measured) + LAPACK getrf/getrs (OpenBLAS) on the %dx%d "
              "leading block%s; %d host threads" % (N, n, n, "" if n == N else " (scaled (N/n)^3)", cores))
    return {"value": (t_asm + t_lu) * 1e3, "unit": "ms", "cores": cores, "kind": "port", "sample": sample,
            "assembly_ms": t_asm * 1e3, "lu_ms" if n == N else "lu_ms_est": t_lu * 1e3}


VENDOR_CHILD = r"""
import ctypes, json, sys, time
import numpy as np
sys.path.insert(0, %(root)r)
import __graft_entry__
fp = __graft_entry__.load_package()
hip = ctypes.CDLL("libamdhip64.so")
rb = ctypes.CDLL("librocblas.so")
rs = ctypes.CDLL("librocsolver.so")
vp = ctypes.c_void_p
hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
hip.hipFree.argtypes = [vp]
rb.rocblas_create_handle.argtypes = [ctypes.POINTER(vp)]
for f in (rs.rocsolver_dgetrf, rs.rocsolver_sgetrf):
    f.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp]
for f in (rs.rocsolver_dgetrs, rs.rocsolver_sgetrs):
    f.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int]
h = vp()
assert rb.rocblas_create_handle(ctypes.byref(h)) == 0

def dmalloc(nbytes):
    p = vp()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p

def run(A, dtype):
    n = A.shape[0]
    A = np.asfortranarray(A, dtype=dtype)
    es = A.itemsize
    dA, dB, dp, di = dmalloc(A.nbytes), dmalloc(n * es), dmalloc(4 * n), dmalloc(4)
    b = np.ones(n, dtype=dtype)
    getrf = rs.rocsolver_dgetrf if dtype == np.float64 else rs.rocsolver_sgetrf
    getrs = rs.rocsolver_dgetrs if dtype == np.float64 else rs.rocsolver_sgetrs
    best = None
    for rep in range(2):
        assert hip.hipMemcpy(dA, A.ctypes.data, A.nbytes, 1) == 0
        assert hip.hipMemcpy(dB, b.ctypes.data, b.nbytes, 1) == 0
        hip.hipDeviceSynchronize()
        t = time.perf_counter()
        assert getrf(h, n, n, dA, n, dp, di) == 0
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        assert getrs(h, 111, n, 1, dA, n, dp, dB, n) == 0
        hip.hipDeviceSynchronize()
        t2 = time.perf_counter()
        if best is None or t1 - t < best[0]:
            best = (t1 - t, t2 - t1)
    info = ctypes.c_int(0)
    hip.hipMemcpy(ctypes.addressof(info), di, 4, 2)
    x = np.empty(n, dtype=dtype)
    hip.hipMemcpy(x.ctypes.data, dB, x.nbytes, 2)
    for p in (dA, dB, dp, di):
        hip.hipFree(p)
    return {"getrf_ms": round(best[0] * 1e3, 2), "getrs_ms": round(best[1] * 1e3, 2), "info": info.value,
            "getrf_tflops": round(2.0 / 3.0 * n ** 3 / best[0] / 1e12, 2)}, x

out = {"what": "rocSOLVER getrf + getrs (1 RHS), matrix resident in HBM, best of 2; a separate process"}
for name, nc, ns, kind in ((%(bwb)r, %(nc)d, %(ns)d, "bwb"), ("wing4800", 60, 80, "wing")):
    if kind == "bwb":
        grid, sh = fp.geometry.bwb(seed=5, n_chord=nc, n_span=ns)
    else:
        grid, sh = fp.geometry.loft_wing(airfoil="0012", n_chord=60, n_span=80, span=10.0, root_chord=1.0,
                                         taper=0.5, sweep_deg=30.0)
    body = fp.RigidWakeBody(grid, sh)
    V = fp.direction(5.0 if kind == "bwb" else 4.2)
    fp.solvers.set_freestream(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
                              np.tile(V[:, None], (1, body.nsheddings)))
    G = fp.assemble(body, "f64")
    out[name + "_f64"], _ = run(G, np.float64)
    if kind == "wing":
        out[name + "_f32"], _ = run(G, np.float32)
    del G
print(json.dumps(out), flush=True)
"""


def vendor_lu(nc, ns):
    """rocSOLVER FP64 (and FP32 at 4,800) getrf + getrs on the same AICs, in a child process started before
    this process touches the GPU: an external anchor for the device LU."""
    import subprocess
    code = VENDOR_CHILD % {"root": ROOT, "bwb": "bwb%dx%d" % (nc, ns), "nc": nc, "ns": ns}
    try:
        p = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=240)
        if p.returncode != 0:
            return {"error": (p.stderr or "")[-300:]}
        return json.loads(p.stdout.strip().splitlines()[-1])
    except Exception as e:  # the anchor is optional
        return {"error": str(e)[:200]}


def c2_leg(fp, ctx, reps=3):
    """The 4,800-panel swept wing (NACA 0012, AR 10, sweep 30, taper 0.5, 60 x 80 quads, AoA 4.2) with one
    freestream: FP32 LU (randomized, no pivoting) + FP64 refinement against the plain FP64 LU."""
    g, sh = fp.geometry.loft_wing(airfoil="0012", n_chord=60, n_span=80, span=10.0, root_chord=1.0, taper=0.5,
                                  sweep_deg=30.0)
    wing = fp.RigidWakeBody(g, sh)
    V = fp.direction(4.2)
    U = np.tile(V[:, None], (1, wing.ncells))
    D = np.tile(V[:, None], (1, wing.nsheddings))
    rec = {"workload": "wing NACA 0012, AR 10, sweep 30, taper 0.5, 60 x 80 quads (4,800 panels) + 80 wake strips, "
                       "AoA 4.2, one freestream", "panels": wing.ncells}
    res = {}
    for name, solver, opts in (("lu32_ir", fp.solve_lu32_ir, {"tol": 1e-10}), ("lu64", fp.solve_ludiv, {})):
        fp.solve(wing, U, D, D, solver=solver, solver_optargs=opts)
        walls, st = [], None
        for _ in range(reps):
            ctx.synchronize()
            t = time.perf_counter()
            st = fp.solve(wing, U, D, D, solver=solver, solver_optargs=opts)
            ctx.synchronize()
            walls.append((time.perf_counter() - t) * 1e3)
        fp.calcfield_U(wing, wing)
        res[name] = (min(walls), st, wing.strength.copy(), fp.calcfield_Cp(wing, 1.0))
    ms32, st32, g32, cp32 = res["lu32_ir"]
    ms64, st64, g64, cp64 = res["lu64"]
    flop = 2.0 / 3.0 * wing.ncells ** 3
    rec.update({"lu32_ir_ms": round(ms32, 3), "lu64_ms": round(ms64, 3), "speedup": round(ms64 / ms32, 2),
                "lu32_ir_stats": st32, "lu64_stats": st64,
                "lu32_factor_tflops": round(flop / (st32["ms_factor"] * 1e-3) / 1e12, 2),
                "lu64_factor_tflops": round(flop / (st64["ms_factor"] * 1e-3) / 1e12, 2),
                "cp_err_max": float(np.abs(cp32 - cp64).max()),
                "gamma_rel_err": float(np.abs(g32 - g64).max() / np.abs(g64).max())})
    return rec


def field_leg(fp, ctx, body, U, n=256):
    """Velocity and Cp of the solved body on an n^3 off-body box grid (Uind!, calcfield_Cp formula)."""
    X = body.grid.nodes
    lo, hi = X.min(axis=1), X.max(axis=1)
    pad = 0.25 * (hi - lo)
    axes = [np.linspace(lo[d] - pad[d], hi[d] + pad[d], n) for d in range(3)]
    g = np.meshgrid(*axes, indexing="ij")
    targets = np.asfortranarray(np.stack([a.ravel() for a in g]))
    del g
    out = np.zeros_like(targets)
    log("bench: field on a %d^3 grid (%d targets)" % (n, targets.shape[1]))
    ctx.synchronize()
    ctx.profile(True)
    t = time.perf_counter()
    fp.Uind(body, targets, out)
    ctx.synchronize()
    wall = (time.perf_counter() - t) * 1e3
    launches, dev_ms = ctx.profile_get("field_velocity_f64")
    ctx.profile(False)
    Uinf = U[:, 0]
    Vt = out + Uinf[:, None]
    cp = 1.0 - (Vt * Vt).sum(axis=0) / Uinf.dot(Uinf)
    pairs = float(targets.shape[1]) * body.ncells
    flop = pairs * 180.0 + float(targets.shape[1]) * 2 * body.nsheddings * 244.0
    return {"targets": int(targets.shape[1]), "grid": "%d^3 box around the body, 25%% margin" % n,
            "device_ms": round(dev_ms, 1), "wall_ms_incl_transfers": round(wall, 1), "launches": launches,
            "pairs": pairs, "flop": flop, "tflops": round(flop / (dev_ms * 1e-3) / 1e12, 2),
            "fp64_frac": round(flop / (dev_ms * 1e-3) / (PEAK_FP64_TFLOPS * 1e12), 3),
            "cp_min": float(cp.min()), "cp_max": float(cp.max())}


def batch_legs(fp, ctx, reps=3):
    """Batched workloads: 64 freestreams as 64 RHS, and 16 uneven wings in one captured graph."""
    out = {}
    # (i) the 4,800-panel swept wing at 64 freestreams (alpha U[-5,10], beta U[-4,4], seed 3),
    # fixed-wake convention (wake along +x for every case): one AIC, 64 right-hand sides
    g, sh = fp.geometry.loft_wing(airfoil="0012", n_chord=60, n_span=80, span=10.0, root_chord=1.0, taper=0.5,
                                  sweep_deg=30.0)
    wing = fp.RigidWakeBody(g, sh)
    rng = np.random.Generator(np.random.PCG64(3))
    al, be = rng.uniform(-5, 10, 64), rng.uniform(-4, 4, 64)
    V = np.stack([fp.direction(a, b) for a, b in zip(al, be)], axis=1)
    U = np.repeat(V[:, None, :], wing.ncells, axis=1)
    D = np.tile(np.array([[1.0], [0.0], [0.0]]), (1, wing.nsheddings))
    rec = {}
    for name, solver in (("lu32_ir", fp.solve_lu32_ir), ("lu64", fp.solve_ludiv)):
        fp.solve_freestreams(wing, U, D, D, solver=solver)
        walls = []
        for _ in range(reps):
            ctx.synchronize()
            t = time.perf_counter()
            G, st = fp.solve_freestreams(wing, U, D, D, solver=solver)
            walls.append((time.perf_counter() - t) * 1e3)
        rec[name] = {"ms": round(min(walls), 3), "stats": st}
        rec[name + "_G"] = G
    # 64 independent single solves for comparison (the same mixed solver)
    ctx.synchronize()
    t = time.perf_counter()
    for c in range(64):
        fp.solve(wing, U[:, :, c], D, D, solver=fp.solve_lu32_ir)
    ctx.synchronize()
    seq_ms = (time.perf_counter() - t) * 1e3
    err = float(np.abs(rec["lu32_ir_G"] - rec["lu64_G"]).max() / np.abs(rec["lu64_G"]).max())
    out["freestreams"] = {
        "workload": "wing NACA 0012, AR 10, sweep 30, taper 0.5, 60 x 80 quads (4,800 panels), 64 freestreams "
                    "(alpha U[-5,10], beta U[-4,4], seed 3), wake along +x for all (fixed-wake convention)",
        "lu32_ir_ms": rec["lu32_ir"]["ms"], "lu64_ms": rec["lu64"]["ms"], "sequential_64_solves_ms": round(seq_ms, 3),
        "gamma_rel_err_mixed_vs_fp64": err, "lu32_ir_stats": rec["lu32_ir"]["stats"],
    }
    # (ii) 16 lofted wings, AR U[6,12], sweep U[0,35] deg, N U[1k,12k] (seed 3), one captured graph
    rng = np.random.Generator(np.random.PCG64(3))
    bodies = []
    for _ in range(16):
        ar, sweep, n = rng.uniform(6, 12), rng.uniform(0, 35), int(rng.uniform(1000, 12000))
        nc = max(10, 2 * int(round(np.sqrt(n / ar))))   # panels around the section: even
        ns = max(4, int(round(n / nc)))
        g, sh = fp.geometry.loft_wing(airfoil="0012", n_chord=nc, n_span=ns, span=ar * 0.75, root_chord=1.0,
                                      taper=0.5, sweep_deg=sweep)
        b = fp.RigidWakeBody(g, sh)
        Vb = fp.direction(4.0)
        fp.solvers.set_freestream(b, np.tile(Vb[:, None], (1, b.ncells)), np.tile(Vb[:, None], (1, b.nsheddings)),
                                  np.tile(Vb[:, None], (1, b.nsheddings)))
        bodies.append(b)
    t = time.perf_counter()
    batch = fp.BodyBatch(bodies, solver=fp.solve_lu32_ir, solver_optargs={"residual_mode": 0, "tol": 1e-10})
    ctx.synchronize()
    create_ms = (time.perf_counter() - t) * 1e3
    batch.solve()
    walls, stats = [], None
    for _ in range(reps):
        ctx.synchronize()
        t = time.perf_counter()
        stats = batch.solve()
        walls.append((time.perf_counter() - t) * 1e3)
    gam = [b.strength.copy() for b in bodies]
    # per-body eager solves with the same solver for comparison
    ctx.synchronize()
    t = time.perf_counter()
    for b in bodies:
        Vb = fp.direction(4.0)
        fp.solve(b, np.tile(Vb[:, None], (1, b.ncells)), np.tile(Vb[:, None], (1, b.nsheddings)),
                 np.tile(Vb[:, None], (1, b.nsheddings)), solver=fp.solve_lu32_ir, solver_optargs={"residual_mode": 0,
                                                                                                   "tol": 1e-10})
    ctx.synchronize()
    seq_ms = (time.perf_counter() - t) * 1e3
    err = max(float(np.abs(g0 - b.strength).max() / np.abs(b.strength).max()) for g0, b in zip(gam, bodies))
    out["bodies"] = {
        "workload": "16 lofted NACA 0012 wings, AR U[6,12], sweep U[0,35] deg, N U[1k,12k] (seed 3), "
                    "FP32 AIC + FP32 LU, matrix-free FP64 residuals",
        "panels": [b.ncells for b in bodies], "total_panels": int(sum(b.ncells for b in bodies)),
        "graph_replay_ms": round(min(walls), 3), "graph_device_ms": round(stats["ms_graph"], 3),
        "create_capture_autotune_ms": round(create_ms, 1), "sequential_solves_ms": round(seq_ms, 3),
        "stats": stats, "autotune": batch.autotune_table(), "gamma_rel_err_vs_sequential": err,
    }
    return out


def adjoint_leg(fp, ctx, fd_coords=64, reps=3, block=400):
    """Lift/drag adjoint of the 10,000-panel wing; returns the JSON record."""
    grid, sh = fp.geometry.loft_wing(airfoil="2412", n_chord=100, n_span=100, span=8.0, root_chord=1.0, taper=1.0,
                                     sweep_deg=0.0, twist_root_deg=0.0, twist_tip_deg=-3.0)
    body = fp.RigidWakeBody(grid, sh)
    U = fp.direction(4.0)
    kw = dict(probe_offset=0.02, probe_length=1.0, span_dir=(0, 1, 0), ref_area=8.0)
    mixed = dict(solver=fp.solve_gmres_ir, solver_optargs={"precond_block": block, "coarse_block": block, "tol": 1e-12,
                                                           "inner_tol": 1e-5, "restart": 100})
    log("bench: adjoint warmup (%d panels)" % body.ncells)
    fp.lift_drag_gradient(body, U, **kw, **mixed)
    ctx.profile(True)
    walls, recs = [], []
    for _ in range(reps):
        ctx.synchronize()
        t = time.perf_counter()
        r = fp.lift_drag_gradient(body, U, **kw, **mixed)
        ctx.synchronize()
        walls.append((time.perf_counter() - t) * 1e3)
        recs.append(r)
        log("bench: adjoint %.1f ms (forward %.1f, reverse %.1f, transpose solve %.1f)"
            % (walls[-1], r["stats"]["ms_forward"], r["stats"]["ms_adjoint"], r["stats"]["ms_transpose_solve"]))
    prof = {k: ctx.profile_get(k) for k in ("adjoint_pairs_probe_f64", "adjoint_pairs_cp_f64", "adjoint_wake_f64",
                                           "transpose", "gemv_f32", "gemv_f64", "assemble_vr_f64",
                                           "field_velocity_f64")}
    ctx.profile(False)
    best = int(np.argmin(walls))
    r = recs[best]
    # FP64-LU adjoint of the same objective (one run) for the mixed/FP64 gradient difference
    r64 = fp.lift_drag_gradient(body, U, **kw)
    g_m = np.concatenate([r["dCL_dnodes"][0].ravel(), r["dCDi_dnodes"][0].ravel()])
    g_64 = np.concatenate([r64["dCL_dnodes"][0].ravel(), r64["dCDi_dnodes"][0].ravel()])
    kernels = {k: {"launches": v[0], "total_ms": round(v[1] / reps, 3)} for k, v in prof.items()}
    rec = {
        "workload": "wing NACA 2412, AR 8, twist 0 -> -3 deg, 100 x 100 vortex-ring quads + rigid wake, AoA 4.0",
        "panels": body.ncells, "nodes": body.grid.nnodes,
        "adjoint_total_ms": round(walls[best], 3),
        "forward_ms": round(r["stats"]["ms_forward"], 3), "reverse_ms": round(r["stats"]["ms_adjoint"], 3),
        "transpose_solve_ms": round(r["stats"]["ms_transpose_solve"], 3),
        "forward_solve": r["stats"]["forward"], "adjoint_iterations": r["stats"]["iterations_adjoint"],
        "CL": r["CL"], "CDi": r["CDi"],
        "mixed_vs_fp64_grad_rel": float(np.linalg.norm(g_m - g_64) / np.linalg.norm(g_64)),
        "fp64_lu_adjoint_total_ms": round(r64["stats"]["ms_forward"] + r64["stats"]["ms_adjoint"], 3),
        "kernels_per_run": kernels,
    }
    if fd_coords <= 0:
        return rec
    # central finite differences of the FP64-LU forward
    log("bench: adjoint FD check (%d node coordinates + Uinf)" % fd_coords)
    rng = np.random.Generator(np.random.PCG64(4))
    idx = rng.choice(3 * body.grid.nnodes, fd_coords, replace=False)
    X0 = np.array(body.grid.nodes)
    h = 1e-6
    fd = np.zeros((2, fd_coords))
    for c, k in enumerate(idx):
        v, d = divmod(int(k), 3)
        vals = []
        for sgn in (1, -1):
            X = X0.copy()
            X[d, v] += sgn * h
            fp.set_nodes(body, X)
            cl, cd, _ = fp.lift_drag(body, U, **kw)
            vals.append((cl, cd))
        fd[:, c] = [(vals[0][0] - vals[1][0]) / (2 * h), (vals[0][1] - vals[1][1]) / (2 * h)]
    fp.set_nodes(body, X0)
    fdu = np.zeros((2, 3))
    for d in range(3):
        vals = []
        for sgn in (1, -1):
            Up = U.copy()
            Up[d] += sgn * h
            cl, cd, _ = fp.lift_drag(body, Up, **kw)
            vals.append((cl, cd))
        fdu[:, d] = [(vals[0][0] - vals[1][0]) / (2 * h), (vals[0][1] - vals[1][1]) / (2 * h)]
    adj = np.stack([r["dCL_dnodes"][0].T.ravel()[idx], r["dCDi_dnodes"][0].T.ravel()[idx]])
    adju = np.stack([r["dCL_dUinf"], r["dCDi_dUinf"]])

    def nrel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    rec["fd_coords"] = int(fd_coords)
    rec["fd_rel_err_nodes"] = {"CL": nrel(adj[0], fd[0]), "CDi": nrel(adj[1], fd[1])}
    rec["fd_rel_err_uinf"] = {"CL": nrel(adju[0], fdu[0]), "CDi": nrel(adju[1], fdu[1])}
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chord", type=int, default=250)
    ap.add_argument("--span", type=int, default=120)
    ap.add_argument("--block", type=int, default=-1, help="block-Jacobi block size (default: four chordwise strips)")
    ap.add_argument("--coarse", type=int, default=-1,
                    help="coarse chunk (default: a twenty-fifth of a chordwise strip)")
    ap.add_argument("--tol", type=float, default=1e-8,
                    help="FP64 relative residual target of the mixed solve (max |dCp| ~1e-8 against FP64 LU, "
                         "100x inside the 1e-6 bound)")
    ap.add_argument("--inner-tol", type=float, default=1e-6, help="relative tolerance of each inner FP32 GMRES")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip the rocSOLVER getrf/getrs anchor on the same AICs (a child process started first)")
    ap.add_argument("--no-c2", action="store_true", help="skip the 4,800-panel single-freestream LU leg")
    ap.add_argument("--cpu-sample", action="store_true",
                    help="CPU baseline LU on a 10,000 x 10,000 block scaled by (N/n)^3 instead of the full 30,000")
    ap.add_argument("--mixed-only", action="store_true", help="profiling runs: mixed steps only (no FP64 LU, no CPU)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="profiling runs: also run a plain 1 GiB device copy to calibrate HBM counters")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 10,000-panel adjoint leg")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg (prints its record)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched legs (64 freestreams, 16 wings)")
    ap.add_argument("--adjoint-fd", type=int, default=64, help="node coordinates checked by finite differences")
    ap.add_argument("--field", type=int, default=256, help="off-body field grid points per axis (0: skip)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched legs (prints their record)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    vendor = None
    if not args.no_vendor and not (args.adjoint_only or args.batch_only or args.mixed_only):
        log("bench: rocSOLVER anchor (child process)")
        vendor = vendor_lu(args.chord, args.span)
    import __graft_entry__
    fp = __graft_entry__.load_package()
    ctx = fp.context()
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(fp, ctx, fd_coords=args.adjoint_fd)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_legs(fp, ctx)}), flush=True)
        return
    alpha = 5.0
    log("bench: building BWB %d x %d" % (args.chord, args.span))
    body, U, D = build_body(fp, args.chord, args.span, alpha)
    N = body.ncells
    fp.solvers.set_freestream(body, U, D, D)
    block = 4 * args.chord if args.block < 0 else args.block
    coarse = max(1, args.chord // 25) if args.coarse < 0 else args.coarse
    mixed_opts = {"precond_block": block, "coarse_block": coarse, "tol": args.tol, "inner_tol": args.inner_tol,
                  "restart": 100, "residual_mode": 1}

    def run(solver, opts):
        return fp.solve(body, U, D, D, solver=solver, solver_optargs=opts)

    for _ in range(args.warmup):
        log("bench: warmup mixed")
        run(fp.solve_gmres_ir, mixed_opts)
    ctx.synchronize()
    ctx.profile(True, only="gemv_f32")   # events around the dominant kernel only
    t0 = time.perf_counter()
    stats = []
    for k in range(args.steps):
        stats.append(run(fp.solve_gmres_ir, mixed_opts))
        log("bench: mixed step %d: %s" % (k, stats[-1]))
    ctx.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps * 1e3
    gemv_timed = ctx.profile_get("gemv_f32")
    ctx.profile(False)
    # per-kernel breakdown of one more (untimed) mixed solve with every scope recording
    ctx.profile(True)
    run(fp.solve_gmres_ir, mixed_opts)
    ctx.synchronize()
    prof = {k: ctx.profile_get(k) for k in ("gemv_f32", "gemv_f64", "assemble_vre_f64", "assemble_vr_f64",
                                           "assemble_wake", "bj_invert", "bj_apply", "coarse_apply")}
    ctx.profile(False)
    gam_mixed = body.strength.copy()
    Us_m = fp.calcfield_U(body, body)
    vs_oracle = oracle_fixture_errors(fp, body, gam_mixed)
    cp_m = fp.calcfield_Cp(body, 1.0)
    if args.copy_calib:
        ctx.check(fp._native.lib().fp_ctx_copy_calibrate(ctx.handle, 1 << 30))
        log("bench: copy calibration: 1 GiB read + 1 GiB written (__amd_rocclr_copyBuffer)")
    if args.mixed_only:
        launches, ms = gemv_timed
        log("bench: gemv_f32 %d launches avg %.4f ms" % (launches, ms / max(launches, 1)))
        print(json.dumps({"mixed_ms": round(t_mixed, 3), "stats": stats[-1], "vs_oracle": vs_oracle}), flush=True)
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
    prof64 = {k: ctx.profile_get(k) for k in ("assemble_vre_f64", "assemble_vr_f64", "gemm_f64", "panel_tslu", "trsm",
                                             "getrs")}
    ctx.profile(False)
    gam64 = body.strength.copy()
    fp.calcfield_U(body, body)
    cp64 = fp.calcfield_Cp(body, 1.0)
    cp_err = float(np.abs(cp_m - cp64).max())
    gam_err = float(np.abs(gam_mixed - gam64).max() / np.abs(gam64).max())
    # the same mixed solve driven to a 1e-10 residual, for reference (one timed run after one warm run)
    strict_opts = dict(mixed_opts, tol=1e-10, inner_tol=1e-5)
    run(fp.solve_gmres_ir, strict_opts)
    ctx.synchronize()
    t = time.perf_counter()
    st_strict = run(fp.solve_gmres_ir, strict_opts)
    ctx.synchronize()
    t_strict = (time.perf_counter() - t) * 1e3
    fp.calcfield_U(body, body)
    strict = {"ms": round(t_strict, 3), "tol": 1e-10, "inner_tol": 1e-5, "stats": st_strict,
              "cp_err_max": float(np.abs(fp.calcfield_Cp(body, 1.0) - cp64).max()),
              "speedup_vs_fp64": round(t_fp64 / t_strict, 3)}
    log("bench: mixed at tol 1e-10: %.1f ms" % t_strict)
    # the mixed solve holding only the FP32 AIC (3.6 GB): FP64 residuals recomputed matrix-free (edge trails)
    mf_opts = dict(mixed_opts, residual_mode=0)
    run(fp.solve_gmres_ir, mf_opts)
    ctx.synchronize()
    ctx.profile(True)
    t = time.perf_counter()
    st_mf = run(fp.solve_gmres_ir, mf_opts)
    ctx.synchronize()
    t_mf = (time.perf_counter() - t) * 1e3
    mf_prof = {k: ctx.profile_get(k) for k in ("field_matvec_f64", "assemble_vre_f64", "gemv_f32")}
    ctx.profile(False)
    fp.calcfield_U(body, body)
    matrix_free = {"ms": round(t_mf, 3), "stats": st_mf, "kernels": mf_prof, "aic_bytes": 4 * N * N,
                   "cp_err_max": float(np.abs(fp.calcfield_Cp(body, 1.0) - cp64).max()),
                   "speedup_vs_fp64": round(t_fp64 / t_mf, 3)}
    log("bench: mixed with matrix-free FP64 residuals: %.1f ms" % t_mf)
    field = field_leg(fp, ctx, body, U, args.field) if args.field > 0 else None

    # roofline of the dominant kernel of the mixed step (FP32 GEMV of the Krylov matvec), timed in the timed steps
    launches, ms = gemv_timed
    avg_ms = ms / max(launches, 1)
    gemv_bytes = 4.0 * N * N + 16.0 * N
    achieved = gemv_bytes / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    roof = {"bound": "hbm", "kernel": "k_gemv<float>", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": GEMV_TRAFFIC_30K if N == 30000 else None,
            "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "algorithmic_bytes_per_launch": gemv_bytes}
    kernels = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in prof.items()}   # one solve
    kernels_fp64 = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in prof64.items()}

    c2 = None if args.no_c2 else c2_leg(fp, ctx)
    adj = None
    if not args.no_adjoint:
        adj = adjoint_leg(fp, ctx, fd_coords=args.adjoint_fd)
    bat = None
    if not args.no_batch:
        bat = batch_legs(fp, ctx)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(fp, body, U, D, full=not args.cpu_sample)

    line = {
        "metric": "AIC+solve time at 30k panels, mixed+refine vs FP64 @ Cp err<=1e-6; adjoint ms",
        "value": round(t_mixed, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mixed, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32 AIC + f64 refinement", "data": "synthetic",
        "config": {"workload": "bwb-%dx%d vortex-ring + rigid wake, AoA %.1f, GMRES-IR (FP32 AIC, block-Jacobi %d, "
                               "coarse chunks %d)" % (args.chord, args.span, alpha, block, coarse),
                   "panels": N, "shedding_edges": body.nsheddings},
        "fp64_lu_ms": round(t_fp64, 3), "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
        "cp_err_max": cp_err, "gamma_rel_err": gam_err, "vs_oracle": vs_oracle, "tol": args.tol, "inner_tol": args.inner_tol,
        "mixed_stats": stats[-1], "fp64_stats": st64, "mixed_tol_1e-10": strict, "mixed_matrix_free": matrix_free,
        "adjoint_ms": adj["adjoint_total_ms"] if adj else None, "adjoint": adj, "c2_wing4800": c2, "batch": bat,
        "field": field,
        "kernels": kernels, "kernels_fp64": kernels_fp64, "vendor_rocsolver": vendor, "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
