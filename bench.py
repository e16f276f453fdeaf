"""Benchmark: PCG solve time to a relative residual of 1e-8 on the
4.36M-unknown vessel network (160^3 capillary lattice + two depth-17
trees, HMIS coarsening), FP32 V-cycle inside flexible FP64 PCG, against
the plain FP64 GPU path and the reference C solve on the host.

A "step" is one complete PCG solve from x0 = 0 with b, the hierarchy and
all work vectors resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                  [--no-cpu] [--no-graph]

--no-graph enqueues every PCG step kernel by kernel instead of replaying
the captured step graph; use it under kernel-tracing profilers.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG time to 1e-8 rel. residual, 4M-node net: FP32-AMG vs FP64 GPU; HBM GB/s"
HBM_PEAK_GBS = 8000.0
TRAFFIC_IC_JACOBI = {"lattice160_trees17": 237249095}  # measured HBM bytes per launch
FAST_COLOURS = 16    # GS on levels with <= 16 colours, l1-Jacobi above
FAST_IC_SWEEPS = 4   # Jacobi sweeps per IC(0) triangle


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(net, coarsening, rtol, maxiter):
    """The reference algorithm's solve on the host, one core, same system.
    PMIS/RS hierarchies use the compiled reference (oracle/_ref, built from
    the C_amg1 sources with gcc -O3).  HMIS uses the oracle restatement
    (gcc -O3, same algorithm): on the 160^3 network the reference's HMIS
    greedy pass raises a node's lambda past the bucket array it allocates
    (C_amg1/src/rs_classical.c:410-418) and writes out of bounds, so the
    compiled reference cannot build this hierarchy.  Sample: the complete
    PCG solve; set-up is not part of the value."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import oracle, reference
    hx = reference() if coarsening != "hmis" else None
    kind = "reference"
    if hx is None:
        hx, kind = oracle(), "port"
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    t0 = time.perf_counter()
    h = hx.build(net.ia, net.ja, net.val, coarsening=coarsening)
    t_setup = time.perf_counter() - t0
    x, it, sec = h.solve(net.b, rtol, maxiter)
    import numpy as np
    import scipy.sparse as sp
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    relres = float(np.linalg.norm(net.b - A @ x) / np.linalg.norm(net.b))
    return {"value": round(sec * 1e3, 3), "unit": "ms", "cores": 1, "kind": kind,
            "sample": (f"complete {'reference' if kind == 'reference' else 'oracle (restated reference)'} "
                       f"PCG solve (AMG+IC, B-norm stop {rtol:g}) on the same "
                       f"{net.n}-unknown system, 1 core of {cpu_model()}; {it} iterations, true "
                       f"relres {relres:.2e}; set-up {t_setup:.1f} s not included")}, x, it


def batch_bench(args, k=32, workload="lattice100_16x16trees11"):
    """Block PCG over k right-hand sides (one solve, CSR SpMM) against a
    loop of k single-right-hand-side solves of the same systems, FP32
    V-cycle in flexible FP64 PCG, relres 1e-8, all data resident in HBM."""
    import numpy as np

    from blood_network_amd import AMGSolver, WORKLOADS, make_network
    from blood_network_amd import _lib as L

    Ledge, depth, ntrees, coarsening = WORKLOADS[workload]
    net = make_network(Ledge, depth, ntrees)
    log(f"[batch] network {workload}: n={net.n} nnz={net.nnz}, {2 * ntrees} Dirichlet roots")
    rng = np.random.default_rng(0)
    art = net.dirichlet_is_art.astype(bool)
    B = np.empty((k, net.n))
    for j in range(k):
        pd = np.where(art, rng.uniform(8e3, 16e3, art.size), rng.uniform(0.5e3, 2e3, art.size))
        B[j] = net.rhs_for(pd)
    t0 = time.perf_counter()
    s = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening, precision="fp32", fine_ic="jacobi",
                  ic_sweeps=FAST_IC_SWEEPS, max_colours=FAST_COLOURS, stop="relres")
    log(f"[batch] hierarchy + upload {time.perf_counter() - t0:.1f} s, {s.num_levels()} levels")
    dB = L.DeviceArray(B)
    dX = L.DeviceArray(np.zeros_like(B))
    zeros = np.zeros_like(B)
    block_ms = []
    for rep in range(args.warmup + args.steps):
        dX.upload(zeros)
        its = s.solve_block_device(k, dB.ptr, dX.ptr, tol=args.rtol)
        if rep >= args.warmup:
            block_ms.append(s.report.solve_ms)
        log(f"[batch] block solve: {s.report.solve_ms:.2f} ms, iterations {int(its.min())}..{int(its.max())}")
    Xb = dX.numpy().reshape(k, net.n)
    loop_ms, loop_its = [], []
    n = net.n
    for rep in range(1 + 1):
        dX.upload(zeros)
        tot, it_list = 0.0, []
        for j in range(k):
            it_list.append(s.solve_device(dB.ptr + 8 * j * n, dX.ptr + 8 * j * n, tol=args.rtol))
            tot += s.report.solve_ms
        if rep >= 1:
            loop_ms.append(tot)
            loop_its = it_list
        log(f"[batch] loop of {k} single solves: {tot:.2f} ms")
    Xl = dX.numpy().reshape(k, n)
    import scipy.sparse as sp
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(n, n))
    relres = np.linalg.norm(B.T - A @ Xb.T, axis=0) / np.linalg.norm(B, axis=1)
    diff = np.linalg.norm(Xb - Xl, axis=1) / np.linalg.norm(Xl, axis=1)
    bm, lm = float(np.median(block_ms)), float(np.median(loop_ms))
    out = {
        "metric": "block PCG throughput, k=32 right-hand sides, 1M-unknown network (solves/s)",
        "value": round(k / (bm * 1e-3), 2), "unit": "solves/s", "higher_is_better": True,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "data": "synthetic (deterministic generator, seed 0)",
        "dtype": "fp32 V-cycle / fp64 PCG",
        "config": {"workload": workload, "n": n, "nnz": net.nnz, "k": k, "coarsening": coarsening,
                   "block_ms": round(bm, 3), "loop_ms": round(lm, 3), "speedup_vs_loop": round(lm / bm, 2),
                   "loop_solves_per_s": round(k / (lm * 1e-3), 2),
                   "block_iterations_max": int(its.max()), "loop_iterations_max": int(max(loop_its)),
                   "relres_max": float(relres.max()), "block_vs_loop_rel_diff_max": float(diff.max())},
    }
    print(json.dumps(out), flush=True)


def networks_bench(args, count=64):
    """64 independent uneven networks: one segmented solve of their
    block-diagonal union (one hierarchy, per-network PCG scalars) against a
    loop of single solves, each network with its own solver.  FP32 V-cycle
    in flexible FP64 PCG, relres 1e-8, data resident in HBM."""
    import numpy as np

    from blood_network_amd import AMGSolver, uneven_batch, union_networks
    from blood_network_amd import _lib as L

    t0 = time.perf_counter()
    nets = uneven_batch(count, seed=0)
    ia, ja, val, b, seg = union_networks(nets)
    n = int(seg[-1])
    log(f"[networks] {count} networks, sizes {min(x.n for x in nets)}..{max(x.n for x in nets)}, "
        f"union n={n} nnz={len(ja)} ({time.perf_counter() - t0:.1f} s)")
    kw = dict(precision="fp32", fine_ic="jacobi", ic_sweeps=FAST_IC_SWEEPS, max_colours=FAST_COLOURS,
              stop="relres")
    t0 = time.perf_counter()
    s = AMGSolver(ia, ja, val, coarsening="pmis", **kw)
    log(f"[networks] union hierarchy + upload {time.perf_counter() - t0:.1f} s, {s.num_levels()} levels")
    db, dx = L.DeviceArray(b), L.DeviceArray(np.zeros(n))
    zeros = np.zeros(n)
    seg_ms = []
    for rep in range(args.warmup + args.steps):
        dx.upload(zeros)
        its = s.solve_segments_device(seg, db.ptr, dx.ptr, tol=args.rtol)
        if rep >= args.warmup:
            seg_ms.append(s.report.solve_ms)
        log(f"[networks] segmented solve: {s.report.solve_ms:.2f} ms, iterations {int(its.min())}..{int(its.max())}")
    xs = dx.numpy()
    # one solver per network, solved one after another
    t0 = time.perf_counter()
    singles = [AMGSolver(net.ia, net.ja, net.val, coarsening="pmis", **kw) for net in nets]
    log(f"[networks] {count} single hierarchies + uploads {time.perf_counter() - t0:.1f} s")
    loop_ms, loop_its = [], []
    for rep in range(2):
        dx.upload(zeros)
        tot, il = 0.0, []
        for g, sg in enumerate(singles):
            il.append(sg.solve_device(db.ptr + 8 * int(seg[g]), dx.ptr + 8 * int(seg[g]), tol=args.rtol))
            tot += sg.report.solve_ms
        if rep >= 1:
            loop_ms.append(tot)
            loop_its = il
        log(f"[networks] loop of {count} single solves: {tot:.2f} ms")
    xl = dx.numpy()
    import scipy.sparse as sp
    A = sp.csr_matrix((val, ja, ia), shape=(n, n))
    r = b - A @ xs
    relres = [np.linalg.norm(r[seg[g]:seg[g + 1]]) / np.linalg.norm(b[seg[g]:seg[g + 1]]) for g in range(count)]
    diff = [np.linalg.norm(xs[seg[g]:seg[g + 1]] - xl[seg[g]:seg[g + 1]]) / np.linalg.norm(xl[seg[g]:seg[g + 1]])
            for g in range(count)]
    sm, lm = float(np.median(seg_ms)), float(np.median(loop_ms))
    out = {
        "metric": "64 uneven networks in one segmented solve (solves/s)",
        "value": round(count / (sm * 1e-3), 2), "unit": "solves/s", "higher_is_better": True,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "data": "synthetic (deterministic generator, seed 0)",
        "dtype": "fp32 V-cycle / fp64 PCG",
        "config": {"networks": count, "lattice_edge": [12, 40], "tree_depth": [6, 10], "union_n": n,
                   "union_nnz": int(len(ja)), "segmented_ms": round(sm, 3), "loop_ms": round(lm, 3),
                   "speedup_vs_loop": round(lm / sm, 2), "loop_solves_per_s": round(count / (lm * 1e-3), 2),
                   "iterations_max": int(its.max()), "loop_iterations_max": int(max(loop_its)),
                   "relres_max": float(max(relres)), "segmented_vs_loop_rel_diff_max": float(max(diff))},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lattice160_trees17")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--kernels-only", action="store_true",
                    help="build the hierarchy, time the fine-level kernels and the copy "
                         "calibration kernel, print them as one JSON line (for counter runs)")
    ap.add_argument("--batch", action="store_true",
                    help="measure the block solve instead: k=32 right-hand sides on the 1M-unknown "
                         "network with 16+16 depth-11 trees, block PCG vs a loop of single solves "
                         "(prints its own JSON line)")
    ap.add_argument("--networks", action="store_true",
                    help="measure 64 uneven networks (lattice edge U{12..40}, tree depth U{6..10}) solved "
                         "in one segmented solve vs a loop of single solves (prints its own JSON line)")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue each PCG step kernel by kernel instead of replaying the captured "
                         "hipGraph (for kernel-tracing profilers)")
    args = ap.parse_args()
    if args.no_graph:
        os.environ["BN_NO_GRAPH"] = "1"
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of this solve is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import numpy as np

    from blood_network_amd import AMGSolver, WORKLOADS, make_network
    from blood_network_amd import _lib as L

    if args.batch:
        return batch_bench(args)
    if args.networks:
        return networks_bench(args)

    Ledge, depth, ntrees, coarsening = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    net = make_network(Ledge, depth, ntrees)
    log(f"[bench] network {args.workload}: n={net.n} nnz={net.nnz} ({time.perf_counter() - t0:.1f} s)")

    # the MI355X path: FP32 V-cycle, Jacobi-swept fine IC(0), flexible FP64
    # PCG; multicolour Gauss-Seidel on levels with at most FAST_COLOURS
    # colours (the others smooth l1-Jacobi)
    t0 = time.perf_counter()
    solver = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening, precision="fp32",
                       fine_ic="jacobi", ic_sweeps=FAST_IC_SWEEPS, max_colours=FAST_COLOURS, stop="relres")
    log(f"[bench] hierarchy + upload {time.perf_counter() - t0:.1f} s, {solver.num_levels()} levels")

    def tk(which, reps=20):
        ms, by = C.c_double(), C.c_double()
        rc = L.lib.bn_solver_time_kernel(C.c_void_p(solver.h), which, reps, C.byref(ms), C.byref(by))
        if rc:
            raise RuntimeError(L.last_error())
        return ms.value, by.value

    if args.kernels_only:
        names = ("fine_spmv_fp64", "fine_gs_sweep", "fine_residual", "fine_ic", "preconditioner", "copy_fp64",
                 "fine_ic_jacobi_sweep")
        res = {}
        for which, nm in enumerate(names):
            ms, by = tk(which, reps=10)
            res[nm] = {"us": round(ms * 1e3, 2), "bytes": by, "GBps": round(by / ms / 1e6, 1) if by else None}
            log(f"[bench] {nm}: {ms * 1e3:.1f} us")
        print(json.dumps({"workload": args.workload, "n": net.n, "nnz": net.nnz, "kernels": res}), flush=True)
        return

    d_b = L.DeviceArray(net.b)
    d_x = L.DeviceArray(np.zeros(net.n))
    zeros = np.zeros(net.n)

    def run(mode):
        solver.set_mode(**mode)
        d_x.upload(zeros)
        it = solver.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
        return it, solver.report

    fp32 = dict(precision="fp32", fine_ic="jacobi", stop="relres")
    fp64j = dict(precision="fp64", fine_ic="jacobi", stop="relres")

    for _ in range(args.warmup):
        run(fp32)
    L.lib.bn_dev_sync()
    t0 = time.perf_counter()
    dev_ms, its = [], []
    for k in range(args.steps):
        it, rep = run(fp32)
        dev_ms.append(rep.solve_ms)
        its.append(it)
        log(f"[bench] step {k}: {it} iterations, {rep.solve_ms:.2f} ms, relres {rep.relres:.2e}")
    L.lib.bn_dev_sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    x32 = d_x.numpy()

    # the same layout and smoothers in FP64
    run(fp64j)
    it64j, rep64j = run(fp64j)
    solver.set_mode(**fp32)
    # plain FP64 GPU path: the reference algorithm's choices (exact IC(0),
    # fixed-beta CG, multicolour GS wherever <= 32 colours), own hierarchy
    plain = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening, stop="relres")
    for _ in range(2):
        d_x.upload(zeros)
        it64 = plain.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
    rep64 = plain.report
    x64 = d_x.numpy()
    plain.free()
    log(f"[bench] fp64 plain: {it64} its {rep64.solve_ms:.2f} ms; fp64 same layout: {it64j} its "
        f"{rep64j.solve_ms:.2f} ms")

    # roofline of the dominant kernel, timed live on the solver's stream
    spmv_ms, spmv_b = tk(0)
    gs_ms, gs_b = tk(1)
    res_ms, res_b = tk(2)
    icj_ms, icj_b = tk(6)
    ach = icj_b / (icj_ms * 1e-3) / 1e9
    kern = {"fine_ic_jacobi_sweep_fp32": (icj_ms, icj_b), "fine_residual_fp32": (res_ms, res_b),
            "fine_spmv_fp64": (spmv_ms, spmv_b), "fine_gs_sweep_fp32": (gs_ms, gs_b)}
    log("[bench] kernels: " + ", ".join(f"{k} {v[0]*1e3:.1f} us {v[1]/v[0]/1e6:.0f} GB/s"
                                        for k, v in kern.items()))

    A = None
    import scipy.sparse as sp
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    relres32 = float(np.linalg.norm(net.b - A @ x32) / np.linalg.norm(net.b))
    err_vs64 = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (1 core) ...")
        cpu, xref, itref = cpu_baseline(net, coarsening, args.rtol, 1000)
        cpu["iterations"] = itref
        cpu["gpu_fp32_rel_err_vs_reference_x"] = float(np.linalg.norm(x32 - xref) / np.linalg.norm(xref))
        log(f"[bench] CPU: {cpu['value']} ms, {itref} iterations")

    value = float(np.median(dev_ms))
    # HBM bytes per launch of the roofline kernel from rocprofv3 FETCH_SIZE
    # and WRITE_SIZE on this workload, each scaled by the same counter's
    # reading for a plain copy kernel of known size (FETCH x2.0, WRITE
    # x1.0), recorded in profiles/ic_jacobi_sweep_lattice160_trees17.md
    traffic = TRAFFIC_IC_JACOBI.get(args.workload)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32 V-cycle / fp64 PCG",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {
            "workload": args.workload, "n": net.n, "nnz": net.nnz, "coarsening": coarsening,
            "levels": solver.num_levels(), "rtol": args.rtol, "stop": "||r||/||b||",
            "iterations": int(np.median(its)),
            "fp64_plain_ms": round(rep64.solve_ms, 3), "fp64_plain_iterations": it64,
            "fp64_same_layout_ms": round(rep64j.solve_ms, 3), "fp64_same_layout_iterations": it64j,
            "max_colours": FAST_COLOURS, "ic_sweeps": FAST_IC_SWEEPS,
            "speedup_vs_fp64_plain": round(rep64.solve_ms / value, 3),
            "speedup_vs_fp64_same_layout": round(rep64j.solve_ms / value, 3),
            "relres_fp32": relres32, "rel_err_vs_fp64": err_vs64,
            "kernels_us": {k: round(v[0] * 1e3, 2) for k, v in kern.items()},
            "kernels_GBps": {k: round(v[1] / v[0] / 1e6, 1) for k, v in kern.items()},
        },
        "roofline": {"bound": "hbm",
                     "kernel": "k_jacobi_rows<float,1>: one Jacobi sweep of the fine IC(0) triangle",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["speedup_vs_cpu_baseline"] = round(cpu["value"] / value, 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
