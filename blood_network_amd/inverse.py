"""Pressure matching on a vessel network: fit segment radii to observed
node pressures.

Loss L(r) = 1/2 sum_k (p_obs[k] - p(r)[obs[k]])^2 over observed nodes, with
p(r) the solution of the pressure system A(g(r)) p = b(g(r)) and
g = pi r^4 / (8 mu len).  The gradient with respect to every segment's
log-radius takes one forward solve, one adjoint solve A lam = dL/dp through
the same hierarchy (A is symmetric) and the edge-scatter kernel
(dL/dlog r_e = -4 g_e (lam_u - lam_v)(p_u - p_v)); all three run in
libbnamg on device-resident arrays.  The host assembles matrix values from
conductances (bn_netgen_assemble) and moves the log-radii.

Between radius sets the preconditioner is rebuilt on the host
(``refresh="rebuild"``: amg_solver_update, a value-only rebuild of the
hierarchy and the fine IC(0) on the stored splitting and interpolation
pattern, C_amg1/src/build_hierarchy.c:170-242, then an upload), rebuilt
on the device with bit-identical results (``refresh="device"``:
bn_solver_rebuild_values_device), or kept from the first radius set
(``refresh="values"``: only the solve matrix's values change on the device,
bn_solver_set_values_device).
"""
import time

import numpy as np

from . import _lib as L
from .amg import AMGSolver

__all__ = ["PressureFit"]


class PressureFit:
    """Loss, gradient and gradient descent for one network and one set of
    observations.  ``solver_kw`` are AMGSolver keyword arguments."""

    def __init__(self, net, obs, p_obs, *, refresh="rebuild", tol=1e-10, maxiter=1000, **solver_kw):
        if refresh not in ("rebuild", "device", "values"):
            raise ValueError("refresh is 'rebuild', 'device' or 'values'")
        self.net = net
        self.obs = np.ascontiguousarray(obs, dtype=np.int64)
        self.p_obs = np.ascontiguousarray(p_obs, dtype=np.float64)
        if self.obs.shape != self.p_obs.shape:
            raise ValueError("one observed pressure per observed node")
        if len(self.obs) and (self.obs.min() < 0 or self.obs.max() >= net.n):
            raise ValueError("observed nodes must be unknowns of the network")
        self.refresh, self.tol, self.maxiter = refresh, tol, maxiter
        self.solver_kw = solver_kw
        self.solver = None
        n, m = net.n, net.n_edges
        self.d_eu, self.d_ev = L.DeviceArray(net.eu), L.DeviceArray(net.ev)
        self.d_pdir = L.DeviceArray(net.p_dirichlet if net.n_dirichlet else np.zeros(1))
        self.d_g = L.DeviceArray(np.zeros(m))
        self.d_grad = L.DeviceArray(np.zeros(m))
        self.d_b, self.d_p = L.DeviceArray(np.zeros(n)), L.DeviceArray(np.zeros(n))
        self.d_lam, self.d_dLdp = L.DeviceArray(np.zeros(n)), L.DeviceArray(np.zeros(n))
        self.d_val = L.DeviceArray(np.zeros(net.nnz)) if refresh != "rebuild" else None
        self.p = None
        self.log = []  # one dict per evaluation: loss, iterations, times

    # ---- one radius set ----------------------------------------------
    def _set_radii(self, logr):
        t0 = time.perf_counter()
        g = self.net.conductance(np.exp(logr))
        val, b = self.net.assemble(g)
        t1 = time.perf_counter()
        if self.solver is None:
            self.solver = AMGSolver(self.net.ia, self.net.ja, val, **self.solver_kw)
        elif self.refresh == "rebuild":
            self.solver.update(self.net.ia, self.net.ja, val)
        elif self.refresh == "device":
            self.d_val.upload(val)
            self.solver.rebuild_values_device(len(val), self.d_val.ptr)
        else:
            self.d_val.upload(val)
            self.solver.set_values_device(len(val), self.d_val.ptr)
        t2 = time.perf_counter()
        self.d_g.upload(g)
        self.d_b.upload(b)
        return {"assemble_ms": (t1 - t0) * 1e3, "refresh_ms": (t2 - t1) * 1e3}

    def loss(self, logr):
        """L at log-radii ``logr``; the pressures stay on the device for a
        following ``grad()`` (the previous pressures are the initial guess)."""
        return self.forward(self._set_radii(np.asarray(logr, dtype=np.float64)))

    def cold_start(self):
        """Zero initial guesses for the next forward and adjoint solves."""
        zeros = np.zeros(self.net.n)
        self.d_p.upload(zeros)
        self.d_lam.upload(zeros)

    def forward(self, rec=None):
        """Pressures and L for the radii the solver holds."""
        rec = {} if rec is None else rec
        it = self.solver.solve_device(self.d_b.ptr, self.d_p.ptr, tol=self.tol, maxiter=self.maxiter)
        if it < 0:
            raise RuntimeError(f"forward solve did not converge ({-it} iterations)")
        rec["forward_its"], rec["forward_ms"] = it, self.solver.report.solve_ms
        self.p = self.d_p.numpy()
        self._res = self.p[self.obs] - self.p_obs
        rec["loss"] = 0.5 * float(self._res @ self._res)
        self.log.append(rec)
        return rec["loss"]

    def grad(self):
        """dL/dlog r at the last ``loss`` point (adjoint warm-started from the
        previous multipliers)."""
        dLdp = np.zeros(self.net.n)
        np.add.at(dLdp, self.obs, self._res)
        self.d_dLdp.upload(dLdp)
        it = self.solver.adjoint_grad_device(self.d_p.ptr, self.d_dLdp.ptr, self.net.n_edges, self.d_eu.ptr,
                                             self.d_ev.ptr, self.d_g.ptr, self.d_pdir.ptr, self.net.n_dirichlet,
                                             self.d_lam.ptr, None, self.d_grad.ptr, tol=self.tol,
                                             maxiter=self.maxiter)
        if it < 0:
            raise RuntimeError(f"adjoint solve did not converge ({-it} iterations)")
        rec = self.log[-1]
        rec["adjoint_its"], rec["adjoint_ms"] = it, self.solver.report.solve_ms
        return self.d_grad.numpy()

    def value_and_grad(self, logr):
        f = self.loss(logr)
        return f, self.grad()

    # ---- descent -----------------------------------------------------
    def descend(self, logr, steps=20, alpha=0.05, armijo=1e-4, grow=1.5, max_halvings=12):
        """Gradient descent on the log-radii with a backtracking (Armijo)
        step: the first trial moves the largest log-radius by ``alpha``;
        a rejected trial halves the step, an accepted one lets the next
        start ``grow`` times larger.  Returns (log-radii, losses) with
        losses[0] at the start and losses[k] after step k."""
        logr = np.array(logr, dtype=np.float64)
        f, g = self.value_and_grad(logr)
        lr = alpha / max(float(np.abs(g).max()), 1e-300)
        losses = [f]
        self.steps = []
        for k in range(steps):
            t0 = time.perf_counter()
            gg, gmax = float(g @ g), float(np.abs(g).max())
            for tries in range(max_halvings + 1):
                trial = logr - lr * g
                ft = self.loss(trial)
                if ft <= f - armijo * lr * gg:
                    break
                lr *= 0.5
            else:
                raise RuntimeError(f"step {k + 1}: no decrease after {max_halvings} halvings")
            logr, f = trial, ft
            g = self.grad()
            losses.append(f)
            self.steps.append({"step": k + 1, "loss": f, "trials": tries + 1, "lr": lr,
                               "max_dlogr": lr * gmax,
                               "wall_ms": (time.perf_counter() - t0) * 1e3})
            lr *= grow
        return logr, losses

    def free(self):
        if self.solver is not None:
            self.solver.free()
            self.solver = None
