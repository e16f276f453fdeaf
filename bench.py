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
# HBM bytes per launch of the roofline kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE
# over its launches in one solve, each counter scaled by its reading for a
# copy kernel of known size), by (workload, kernel symbol).  lattice160_trees17,
# per launch: the level-1 Gauss-Seidel colours (527 launches per solve) FETCH
# 32.45 MB (counter x 1.999) + WRITE 3.00 MB (x 1.000), against 20.06 MB
# algorithmic (1.77x); the residuals (102 launches) 205.0 + 7.6 MB.
TRAFFIC = {("lattice160_trees17", "k_rowsolve<float, 8, 2>"): 35445339,
           ("lattice160_trees17", "k_csr_blocks<float, 1, false>"): 212617119}
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


def descent_bench(args, n_obs=10_000, steps=20):
    """Radius gradients on the 1,065,532-unknown network (100^3 lattice,
    depth-15 trees): true radii = base x lognormal(0, 0.1), pressures
    observed at 10,000 random nodes, L = 1/2 sum (p_obs - p)^2.  Times the
    gradient (forward solve + adjoint solve through the same hierarchy +
    edge scatter, zero initial guesses, radii fixed) in FP64 and with the
    FP32 V-cycle in flexible FP64 PCG, each against an FP64 gradient solved
    to 1e-13; then 20 descent steps with the hierarchy rebuilt per radius
    set on the host, rebuilt on the device, and kept (device value swap).  The CPU leg: the compiled
    reference's forward and adjoint PCG solves on one host core."""
    import numpy as np

    from blood_network_amd import AMGSolver, PressureFit, make_network

    net = make_network(100, 15, 1, seed=0)
    rng = np.random.default_rng(4)
    logr0 = np.log(net.radius)
    logr_true = logr0 + rng.normal(0.0, 0.1, net.n_edges)
    obs = rng.choice(net.n, size=n_obs, replace=False)
    vt, bt = net.assemble(net.conductance(np.exp(logr_true)))
    st = AMGSolver(net.ia, net.ja, vt, coarsening="pmis")
    p_obs = st.solve(bt, tol=1e-13, maxiter=500)[0][obs]
    st.free()
    ref = PressureFit(net, obs, p_obs, tol=1e-13, coarsening="pmis")
    f_ref, g_ref = ref.value_and_grad(logr0)
    ref.free()
    log(f"[descent] n={net.n} segments={net.n_edges}; loss at base radii {f_ref:.6e}")
    modes = {"fp64": dict(coarsening="pmis"),
             "fp32_flex": dict(coarsening="pmis", precision="fp32", fine_ic="jacobi", ic_sweeps=FAST_IC_SWEEPS,
                               max_colours=FAST_COLOURS)}
    grad = {}
    for name, kw in modes.items():
        fit = PressureFit(net, obs, p_obs, tol=args.rtol, **kw)
        fit.value_and_grad(logr0)  # set-up, upload, graph capture
        wall = []
        for rep in range(args.warmup + args.steps):
            fit.cold_start()
            t0 = time.perf_counter()
            f = fit.forward()
            g = fit.grad()
            if rep >= args.warmup:
                wall.append((time.perf_counter() - t0) * 1e3)
        rec = fit.log[-1]
        err = float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
        grad[name] = {"ms": round(float(np.median(wall)), 3), "forward_its": rec["forward_its"],
                      "adjoint_its": rec["adjoint_its"], "forward_device_ms": round(rec["forward_ms"], 3),
                      "adjoint_device_ms": round(rec["adjoint_ms"], 3), "rel_err_vs_fp64_1e-13": err,
                      "loss_rel_err": abs(f - f_ref) / f_ref}
        log(f"[descent] gradient {name}: {grad[name]}")
        fit.free()
    descents = {}
    for refresh in ("rebuild", "device", "values"):
        fit = PressureFit(net, obs, p_obs, tol=args.rtol, refresh=refresh, **modes["fp32_flex"])
        t0 = time.perf_counter()
        _, losses = fit.descend(logr0, steps=steps)
        total = time.perf_counter() - t0
        evals = fit.log
        descents[refresh] = {
            "losses": [float("%.6e" % x) for x in losses], "seconds": round(total, 2),
            "step_ms_median": round(float(np.median([s["wall_ms"] for s in fit.steps])), 1),
            "refresh_ms_median": round(float(np.median([e["refresh_ms"] for e in evals[1:]])), 1),
            "forward_its": [e["forward_its"] for e in evals],
            "trials": [s["trials"] for s in fit.steps]}
        log(f"[descent] {steps} steps, refresh={refresh}: loss {losses[0]:.4e} -> {losses[-1]:.4e} "
            f"in {total:.1f} s; median step {descents[refresh]['step_ms_median']} ms")
        fit.free()
    # CPU leg last (it pins the process to one core)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import oracle, reference
        hx, kind = reference(), "reference"
        if hx is None:
            hx, kind = oracle(), "port"
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except (AttributeError, OSError):
            pass
        g0 = net.conductance(np.exp(logr0))
        v0, b0 = net.assemble(g0)
        h = hx.build(net.ia, net.ja, v0, coarsening="pmis")
        t0 = time.perf_counter()
        p, it1, _ = h.solve(b0, args.rtol, 1000)
        d = np.zeros(net.n)
        np.add.at(d, obs, p[obs] - p_obs)
        lam, it2, _ = h.solve(d, args.rtol, 1000)
        pd = net.p_dirichlet
        pu = np.where(net.eu >= 0, p[np.maximum(net.eu, 0)], pd[np.maximum(-1 - net.eu, 0)])
        pv = np.where(net.ev >= 0, p[np.maximum(net.ev, 0)], pd[np.maximum(-1 - net.ev, 0)])
        lu = np.where(net.eu >= 0, lam[np.maximum(net.eu, 0)], 0.0)
        lv = np.where(net.ev >= 0, lam[np.maximum(net.ev, 0)], 0.0)
        gc = 4.0 * g0 * (-(lu - lv) * (pu - pv))
        sec = time.perf_counter() - t0
        cpu = {"value": round(sec * 1e3, 1), "unit": "ms", "cores": 1, "kind": kind,
               "sample": (f"forward + adjoint PCG solves (AMG+IC, B-norm stop {args.rtol:g}; {it1} + {it2} "
                          f"iterations) and the numpy edge formula on the same system, 1 core of {cpu_model()}; "
                          f"gradient rel. error vs the FP64 1e-13 GPU gradient "
                          f"{float(np.linalg.norm(gc - g_ref) / np.linalg.norm(g_ref)):.2e}")}
        log(f"[descent] CPU gradient {cpu['value']} ms")
    value = grad["fp32_flex"]["ms"]
    out = {
        "metric": "radius gradient, 1.07M unknowns / 3.07M segments (forward + adjoint + edge scatter)",
        "value": value, "unit": "ms", "higher_is_better": False, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "data": "synthetic (deterministic generator, seed 0; observations seed 4)",
        "dtype": "fp32 V-cycle / fp64 PCG and scatter", "vs_baseline": None,
        "config": {"workload": "lattice100_trees15 radius fit", "unknowns": net.n, "segments": net.n_edges,
                   "observations": n_obs, "rtol_bnorm": args.rtol, "gradient": grad, "descent": descents},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["speedup_vs_cpu_baseline"] = round(cpu["value"] / value, 1)
    print(json.dumps(out), flush=True)


def plain_bench(args, workload="lattice100_trees15"):
    """The plain FP64 method on one GPU against the reference on one host
    core: the 1,065,532-unknown network (100^3 lattice, depth-15 trees),
    PMIS, single right-hand side, PCG to the reference's stop test
    sqrt(r'z / r0'z0) < 1e-8 (C_amg1/src/amg_pcg.c:75).  A step is one
    complete solve from x0 = 0 with every array resident in HBM.  Two GPU
    forms: the drop-in (amg_solver_create's choices: multicolour C-then-F
    Gauss-Seidel, exact level-scheduled IC(0), dense coarsest inverse) is the
    value; the reference's exact algorithm (cf_order sweeps level-scheduled,
    IC-PCG coarsest solve) is reported beside it.  The CPU leg is the compiled
    reference (oracle/_ref, C_amg1 built -O3), whose solution x_ref and
    iteration count the GPU solves are compared with."""
    import numpy as np
    import scipy.sparse as sp

    from blood_network_amd import AMGSolver, WORKLOADS, workload_network
    from blood_network_amd import _lib as L

    Ledge, depth, ntrees, coarsening = WORKLOADS[workload]
    net = workload_network(workload)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    log(f"[plain] network {workload}: n={net.n} nnz={net.nnz}")
    t0 = time.perf_counter()
    s = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening)
    log(f"[plain] hierarchy + upload {time.perf_counter() - t0:.1f} s, {s.num_levels()} levels")
    d_b = L.DeviceArray(net.b)
    d_x = L.DeviceArray(np.zeros(net.n))
    zeros = np.zeros(net.n)
    for _ in range(args.warmup):
        d_x.upload(zeros)
        s.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
    L.lib.bn_dev_sync()
    t0 = time.perf_counter()
    dev_ms = []
    for k in range(args.steps):
        d_x.upload(zeros)
        it = s.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
        dev_ms.append(s.report.solve_ms)
        log(f"[plain] step {k}: {it} iterations, {s.report.solve_ms:.2f} ms")
    L.lib.bn_dev_sync()
    wall = time.perf_counter() - t0
    x = d_x.numpy()
    s.free()
    # the reference's exact algorithm on the GPU
    ex = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening, coarse="icpcg", gs_exact=True)
    d_x.upload(zeros)
    ex.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
    d_x.upload(zeros)
    it_ex = ex.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
    ms_ex = ex.report.solve_ms
    x_ex = d_x.numpy()
    ex.free()
    log(f"[plain] exact reference order on the GPU: {it_ex} iterations, {ms_ex:.1f} ms")
    cpu, xref, itref = None, None, None
    if not args.no_cpu:
        log("[plain] CPU baseline: the compiled reference, 1 core ...")
        cpu, xref, itref = cpu_baseline(net, coarsening, args.rtol, 1000)
        cpu["iterations"] = itref
    value = float(np.median(dev_ms))

    def err(v):
        return float(np.linalg.norm(v - xref) / np.linalg.norm(xref)) if xref is not None else None

    out = {
        "metric": "plain FP64 PCG solve time to the reference's 1e-8 stop, 1M-unknown network (vs reference on host)",
        "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": workload, "n": net.n, "nnz": net.nnz, "coarsening": coarsening,
                   "rtol_bnorm": args.rtol, "iterations": it,
                   "true_relres": float(np.linalg.norm(net.b - A @ x) / np.linalg.norm(net.b)),
                   "rel_err_vs_reference_x": err(x), "reference_iterations": itref,
                   "exact_reference_order": {"ms": round(ms_ex, 3), "iterations": it_ex,
                                             "rel_err_vs_reference_x": err(x_ex)}},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["speedup_vs_cpu_baseline"] = round(cpu["value"] / value, 2)
    print(json.dumps(out), flush=True)


def batch_bench(args, k=32, workload="lattice100_16x16trees11"):
    """Block PCG over k right-hand sides (one solve, CSR SpMM) against a
    loop of k single-right-hand-side solves of the same systems, FP32
    V-cycle in flexible FP64 PCG, relres 1e-8, all data resident in HBM."""
    import numpy as np

    from blood_network_amd import AMGSolver, WORKLOADS, workload_network
    from blood_network_amd import _lib as L

    Ledge, depth, ntrees, coarsening = WORKLOADS[workload]
    net = workload_network(workload)
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


def plain_headline_bench(args):
    """The plain FP64 GPU path alone on the headline network (4.36M
    unknowns, HMIS): amg_solver_create's choices (multicolour C-then-F GS on
    levels of <= 32 colours, exact IC(0) solved one launch per triangle,
    dense coarsest inverse, fixed-beta CG; C_amg1/src/amg_pcg.c:12-101),
    stopping at ||r||/||b|| < rtol.  A step is one solve from x0 = 0 with
    every array resident in HBM; prints its own JSON line (profiling runs
    of the drop-in path use it)."""
    import numpy as np

    from blood_network_amd import AMGSolver, WORKLOADS, make_network
    from blood_network_amd import _lib as L

    Ledge, depth, ntrees, coarsening = WORKLOADS[args.workload]
    net = make_network(Ledge, depth, ntrees)
    plain = AMGSolver(net.ia, net.ja, net.val, coarsening=coarsening, stop="relres")
    d_b = L.DeviceArray(net.b)
    d_x = L.DeviceArray(np.zeros(net.n))
    zeros = np.zeros(net.n)
    for _ in range(args.warmup):
        d_x.upload(zeros)
        plain.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
    ms = []
    for k in range(args.steps):
        d_x.upload(zeros)
        it = plain.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000)
        ms.append(plain.report.solve_ms)
        log(f"[plain] step {k}: {it} iterations, {ms[-1]:.2f} ms")
    plain.free()
    print(json.dumps({
        "metric": "plain FP64 PCG solve time to 1e-8 rel. residual, 4M-node net (drop-in path)",
        "value": round(float(np.median(ms)), 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "fp64",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": args.workload, "n": net.n, "nnz": net.nnz, "coarsening": coarsening,
                   "rtol": args.rtol, "iterations": it}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lattice160_trees17")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rtol", type=float, default=None,
                    help="PCG tolerance (default 1e-8; 1e-10 B-norm for --descent)")
    ap.add_argument("--counters", action="store_true",
                    help="for counter runs: build the hierarchy, run --steps solves on the benchmark's path "
                         "with every kernel enqueued eagerly, then 5 plain FP64 copies of n values (the "
                         "byte-counter calibration), print one JSON line")
    ap.add_argument("--kernels-only", action="store_true",
                    help="build the hierarchy, time the fine-level kernels and the copy "
                         "calibration kernel, print them as one JSON line (for counter runs)")
    ap.add_argument("--batch", action="store_true",
                    help="measure the block solve instead: k=32 right-hand sides on the 1M-unknown "
                         "network with 16+16 depth-11 trees, block PCG vs a loop of single solves "
                         "(prints its own JSON line)")
    ap.add_argument("--descent", action="store_true",
                    help="radius gradients and 20 descent steps on the 1.07M-unknown network "
                         "(prints its own JSON line)")
    ap.add_argument("--networks", action="store_true",
                    help="measure 64 uneven networks (lattice edge U{12..40}, tree depth U{6..10}) solved "
                         "in one segmented solve vs a loop of single solves (prints its own JSON line)")
    ap.add_argument("--plain", action="store_true",
                    help="time only the plain FP64 path (the drop-in's choices) on the headline network "
                         "(prints its own JSON line)")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue each PCG step kernel by kernel instead of replaying the captured "
                         "hipGraph (for kernel-tracing profilers)")
    args = ap.parse_args()
    if args.rtol is None:
        args.rtol = 1e-10 if args.descent else 1e-8
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
    if args.descent:
        return descent_bench(args)
    if args.workload == "lattice100_trees15":
        return plain_bench(args)
    if args.plain:
        return plain_headline_bench(args)

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

    if args.counters:
        os.environ["BN_NO_GRAPH"] = "1"
        d_b = L.DeviceArray(net.b)
        d_x = L.DeviceArray(np.zeros(net.n))
        its = []
        for _ in range(args.steps):
            d_x.upload(np.zeros(net.n))
            its.append(solver.solve_device(d_b.ptr, d_x.ptr, tol=args.rtol, maxiter=1000))
        copy_ms, copy_b = tk(5, reps=5)
        print(json.dumps({"workload": args.workload, "n": net.n, "solves": args.steps, "iterations": its,
                          "copy_bytes": copy_b, "copy_us": round(copy_ms * 1e3, 2)}), flush=True)
        return

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
    # the FP64 reference solution the FP32 path is measured against: the
    # plain FP64 GPU solve to ||r||/||b|| < 1e-12
    d_x.upload(zeros)
    it64_12 = plain.solve_device(d_b.ptr, d_x.ptr, tol=1e-12, maxiter=1000)
    x64 = d_x.numpy()
    plain.free()
    log(f"[bench] fp64 plain: {it64} its {rep64.solve_ms:.2f} ms; fp64 same layout: {it64j} its "
        f"{rep64j.solve_ms:.2f} ms")

    # kernel census of one more solve on the benchmark's path: every launch
    # of its bandwidth kernels timed by HIP events on the solver's stream as
    # the solve runs (and each launch shape again cold); the roofline kernel
    # is the symbol with the most device time per solve
    solver.set_mode(**fp32)
    d_x.upload(zeros)
    census = solver.kernel_census(d_b.ptr, d_x.ptr, tol=args.rtol, reps=3)
    top = census[0]
    ach = top["bytes"] / (top["us_solve"] * 1e-6) / 1e9
    ach_cold = top["bytes"] / (top["us_cold"] * 1e-6) / 1e9
    for c in census[:8]:
        log(f"[bench] census {c['name']}: {c['launches']} launches, {c['us_solve'] / 1e3:.3f} ms per solve "
            f"({c['bytes'] / c['us_solve'] / 1e3:.0f} GB/s as run, {c['bytes'] / c['us_cold'] / 1e3:.0f} GB/s cold)")
    # single fine-level kernels (HIP events over repeated launches)
    spmv_ms, spmv_b = tk(0)
    gs_ms, gs_b = tk(1)
    res_ms, res_b = tk(2)
    icj_ms, icj_b = tk(6)
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
    traffic = TRAFFIC.get((args.workload, top["name"]))
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
            "relres_fp32": relres32, "rel_err_vs_fp64_1e-12": err_vs64, "fp64_1e-12_iterations": it64_12,
            "kernels_us": {k: round(v[0] * 1e3, 2) for k, v in kern.items()},
            "kernels_GBps": {k: round(v[1] / v[0] / 1e6, 1) for k, v in kern.items()},
            "kernel_census": [{"kernel": c["name"], "launches": c["launches"], "ms": round(c["us_solve"] / 1e3, 3),
                               "GBps": round(c["bytes"] / c["us_solve"] / 1e3, 1),
                               "GBps_cold": round(c["bytes"] / c["us_cold"] / 1e3, 1)} for c in census[:10]],
        },
        "roofline": {"bound": "hbm", "kernel": top["name"],
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launches_per_solve": top["launches"],
                     "bytes_per_launch": round(top["bytes"] / top["launches"]),
                     "us_per_launch": round(top["us_solve"] / top["launches"], 2),
                     "cold": {"achieved": round(ach_cold, 1), "frac": round(ach_cold / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["speedup_vs_cpu_baseline"] = round(cpu["value"] / value, 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
