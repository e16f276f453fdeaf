"""Radius gradients at full network scale: the 1,065,532-unknown network
(100^3 capillary lattice, arterial and venous trees of depth 15), true
radii = base radii x lognormal(0, 0.1) per segment, pressures observed at
10,000 random nodes, L = 1/2 sum (p_obs - p(r))^2.

The device gradient (forward solve, adjoint solve through the same
hierarchy, edge-scatter kernel) is checked against an FP64 CPU reference
gradient (the oracle's AMG-PCG on the host, numpy edge formula), against
central differences of the loss on 64 random segments (relative radius
step 1e-6, FP64), and then drives 20 gradient-descent steps."""
import sys
import time

import numpy as np
import pytest

from blood_network_amd import AMGSolver, PressureFit, make_network
from blood_network_amd import _lib as L

pytestmark = pytest.mark.gpu

N_OBS = 10_000
FD_EDGES = 64
FD_STEP = 1e-6  # in log r: a relative radius step


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def edge_pressures(net, x):
    """x at both ends of every segment (Dirichlet ends: their pressure)."""
    pd = net.p_dirichlet
    xu = np.where(net.eu >= 0, x[np.maximum(net.eu, 0)], pd[np.maximum(-1 - net.eu, 0)])
    xv = np.where(net.ev >= 0, x[np.maximum(net.ev, 0)], pd[np.maximum(-1 - net.ev, 0)])
    return xu, xv


def edge_values(net, x):
    """x at both ends, zero at Dirichlet ends (multipliers vanish there)."""
    xu = np.where(net.eu >= 0, x[np.maximum(net.eu, 0)], 0.0)
    xv = np.where(net.ev >= 0, x[np.maximum(net.ev, 0)], 0.0)
    return xu, xv


@pytest.fixture(scope="module")
def problem():
    case("network 100^3 lattice + depth-15 trees; true radii; 10,000 observations")
    t0 = time.perf_counter()
    net = make_network(100, 15, 1, seed=0)
    assert net.n == 1_065_532
    rng = np.random.default_rng(4)
    logr0 = np.log(net.radius)
    logr_true = logr0 + rng.normal(0.0, 0.1, net.n_edges)
    obs = rng.choice(net.n, size=N_OBS, replace=False)
    val, b = net.assemble(net.conductance(np.exp(logr_true)))
    s = AMGSolver(net.ia, net.ja, val, coarsening="pmis")
    p_true, it = s.solve(b, tol=1e-13, maxiter=500)
    s.free()
    assert it > 0
    print(f"observations from the true radii: {it} iterations, {time.perf_counter() - t0:.1f} s", flush=True)
    return net, logr0, obs, p_true[obs]


@pytest.fixture(scope="module")
def base_gradient(problem):
    net, logr0, obs, p_obs = problem
    fit = PressureFit(net, obs, p_obs, tol=1e-13, coarsening="pmis")
    t0 = time.perf_counter()
    f, grad = fit.value_and_grad(logr0)
    rec = fit.log[-1]
    print(f"device loss {f:.6e}; forward {rec['forward_its']} its {rec['forward_ms']:.1f} ms, adjoint "
          f"{rec['adjoint_its']} its {rec['adjoint_ms']:.1f} ms ({time.perf_counter() - t0:.1f} s with set-up)",
          flush=True)
    yield fit, f, grad
    fit.free()


def test_gradient_matches_fp64_cpu_reference(problem, base_gradient):
    case("device gradient vs FP64 CPU reference gradient (oracle AMG-PCG on the host)")
    import oracle_lib

    net, logr0, obs, p_obs = problem
    fit, f, grad = base_gradient
    g0 = net.conductance(np.exp(logr0))
    val, b = net.assemble(g0)
    t0 = time.perf_counter()
    h = oracle_lib.oracle().build(net.ia, net.ja, val, coarsening="pmis")
    p_ref = h.solve(b, 1e-13, 500)[0]
    res = p_ref[obs] - p_obs
    f_ref = 0.5 * float(res @ res)
    dLdp = np.zeros(net.n)
    np.add.at(dLdp, obs, res)
    lam_ref = h.solve(dLdp, 1e-13, 500)[0]
    pu, pv = edge_pressures(net, p_ref)
    lu, lv = edge_values(net, lam_ref)
    ref = 4.0 * g0 * (-(lu - lv) * (pu - pv))
    err = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    worst = np.abs(grad - ref).max() / np.abs(ref).max()
    print(f"CPU reference {time.perf_counter() - t0:.1f} s; loss {f:.10e} vs {f_ref:.10e}; "
          f"gradient rel. 2-norm error {err:.2e}, max-entry error / max entry {worst:.2e}", flush=True)
    assert abs(f - f_ref) <= 1e-9 * f_ref
    assert err < 1e-6
    assert worst < 1e-6


def test_gradient_matches_central_differences(problem, base_gradient):
    """L(log r + h e_k) - L(log r - h e_k) on 64 random segments.  Each
    perturbed pressure field is p0 + d, with d from the perturbed system
    A_pm d = (b_pm - b0) - (A_pm - A0) p0, solved to 1e-11 relative to its
    own right-hand side: a segment changes only its own conductance, so
    the right-hand side sits on the segment's two ends, and d carries the
    perturbation at full relative accuracy (a from-scratch solve of p_pm
    would bury a 1e-6 step under the base solve's rounding).  Then
    L_pm - L0 = r0.d + d.d/2 on the observed nodes exactly."""
    case("device gradient vs central differences on 64 random segments")
    net, logr0, obs, p_obs = problem
    fit, f, grad = base_gradient
    s, p0 = fit.solver, fit.p.copy()
    r0 = p0[obs] - p_obs
    g0 = net.conductance(np.exp(logr0))
    val0, _ = net.assemble(g0)
    rng = np.random.default_rng(64)
    edges = rng.choice(net.n_edges, size=FD_EDGES, replace=False)
    d_val = L.DeviceArray(val0)
    d_rhs, d_x = L.DeviceArray(np.zeros(net.n)), L.DeviceArray(np.zeros(net.n))
    zeros = np.zeros(net.n)
    pu_all, pv_all = edge_pressures(net, p0)
    gmax = np.abs(grad).max()
    worst_rel, its = 0.0, []
    t0 = time.perf_counter()
    try:
        for e in edges:
            u, v = int(net.eu[e]), int(net.ev[e])
            dp = pu_all[e] - pv_all[e]
            d = {}
            for sgn in (1, -1):
                g = g0.copy()
                g[e] = g0[e] * np.exp(4.0 * sgn * FD_STEP)
                val, _ = net.assemble(g)
                d_val.upload(val)
                s.set_values_device(len(val), d_val.ptr)
                delta = g[e] - g0[e]
                rhs = np.zeros(net.n)
                if u >= 0:
                    rhs[u] -= delta * dp
                if v >= 0:
                    rhs[v] += delta * dp
                if not rhs.any():
                    d[sgn] = zeros
                    continue
                d_rhs.upload(rhs)
                d_x.upload(zeros)
                it = s.solve_device(d_rhs.ptr, d_x.ptr, tol=1e-11, maxiter=500)
                assert it > 0
                its.append(it)
                d[sgn] = d_x.numpy()
            dpl, dmi = d[1][obs], d[-1][obs]
            fd = (r0 @ (dpl - dmi) + 0.5 * (dpl @ dpl - dmi @ dmi)) / (2 * FD_STEP)
            err = abs(fd - grad[e])
            rel = err / max(abs(fd), 1e-300)
            print(f"  segment {e:8d} ({u:8d},{v:8d}) central difference {fd: .10e} adjoint {grad[e]: .10e} "
                  f"rel {rel:.1e}", flush=True)
            if abs(fd) > 1e-6 * gmax:
                worst_rel = max(worst_rel, rel)
            assert err <= 1e-5 * abs(fd) + 1e-9 * gmax, (e, fd, grad[e])
    finally:
        d_val.upload(val0)
        s.set_values_device(len(val0), d_val.ptr)
    print(f"{FD_EDGES} segments, {time.perf_counter() - t0:.1f} s, correction solves {min(its)}..{max(its)} its; "
          f"worst relative difference (|fd| > 1e-6 max|grad|) {worst_rel:.2e}", flush=True)


def test_twenty_descent_steps(problem):
    """20 gradient steps on the log-radii: the loss decreases every step;
    keeping the first hierarchy (values-only refresh) follows the same
    path as rebuilding it each step for the first steps."""
    case("20 gradient-descent steps, hierarchy rebuilt every radius set")
    net, logr0, obs, p_obs = problem
    fit = PressureFit(net, obs, p_obs, tol=1e-11, coarsening="pmis")
    t0 = time.perf_counter()
    try:
        _, losses = fit.descend(logr0, steps=20)
    finally:
        fit.free()
    for st in fit.steps:
        print(f"  step {st['step']:2d} loss {st['loss']:.6e} trials {st['trials']} "
              f"max |d log r| {st['max_dlogr']:.3e} {st['wall_ms']:.0f} ms", flush=True)
    print(f"loss {losses[0]:.6e} -> {losses[-1]:.6e} ({losses[-1] / losses[0]:.3f}) in "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.9 * losses[0]

    case("5 steps keeping the first hierarchy (device value refresh) follow the same losses")
    fit2 = PressureFit(net, obs, p_obs, tol=1e-11, coarsening="pmis", refresh="values")
    try:
        _, losses2 = fit2.descend(logr0, steps=5)
        its = [(r["forward_its"], r.get("adjoint_its")) for r in fit2.log]
    finally:
        fit2.free()
    print(f"values-only refresh losses {['%.6e' % x for x in losses2]}; iterations {its}", flush=True)
    np.testing.assert_allclose(losses2, losses[:6], rtol=1e-7)

    case("5 steps rebuilding the hierarchy on the device follow the host rebuild's losses")
    fit3 = PressureFit(net, obs, p_obs, tol=1e-11, coarsening="pmis", refresh="device")
    try:
        _, losses3 = fit3.descend(logr0, steps=5)
        refresh_ms = [round(r["refresh_ms"], 1) for r in fit3.log[1:]]
    finally:
        fit3.free()
    print(f"device-rebuild losses {['%.6e' % x for x in losses3]}; refresh ms {refresh_ms}", flush=True)
    np.testing.assert_allclose(losses3, losses[:6], rtol=1e-9)
