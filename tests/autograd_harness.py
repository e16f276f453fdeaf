"""torch.autograd.Function over the pressure solve (test harness only).

Run as its own process (tests/test_gpu_autograd.py starts it): torch is
imported before the package, so libbnamg resolves libamdhip64.so.7 and
libhsa-runtime64.so.1 to the copies torch already loaded and the process
has one HIP runtime; CUDA tensors' data_ptr() values are then valid
device pointers for the library's device-pointer entry points.

PressureSolve maps segment log-radii to node pressures: forward is the
AMG-PCG solve, backward is the adjoint solve through the same hierarchy
plus the edge-scatter kernel, both through the C ABI.  Checks: the
gradient of a pressure-matching loss through torch equals the C ABI's own
adjoint gradient and an FP64 direct-solve gradient (scipy); CUDA tensors
give the CPU-tensor gradient; an optimizer lowers the loss.  Prints one
JSON line and exits 0 when every check holds."""
import json
import os
import sys
import time

import torch  # noqa: I001  (first: one HIP runtime for torch and libbnamg)

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from blood_network_amd import AMGSolver, make_network  # noqa: E402


def say(msg):
    print(msg, flush=True)


class NetworkModel:
    """One network and the solver whose values follow the last radii."""

    def __init__(self, net, tol=1e-13, **solver_kw):
        self.net, self.tol, self.kw = net, tol, solver_kw
        self.solver = None
        self.current = None  # the conductances the solver holds
        self.dev = {}

    def refresh(self, g, val):
        if self.current is g:
            return
        if self.solver is None:
            self.solver = AMGSolver(self.net.ia, self.net.ja, val, **self.kw)
        else:
            self.solver.update(self.net.ia, self.net.ja, val)
        self.current = g

    def device_edges(self, device):
        if device not in self.dev:
            net = self.net
            self.dev[device] = (torch.from_numpy(net.eu).to(device), torch.from_numpy(net.ev).to(device),
                                torch.from_numpy(net.p_dirichlet).to(device))
        return self.dev[device]


class PressureSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, logr, model):
        net = model.net
        g = net.conductance(np.exp(logr.detach().cpu().numpy()))
        val, b = net.assemble(g)
        model.refresh(g, val)
        if logr.is_cuda:
            bt = torch.from_numpy(b).to(logr.device)
            p = torch.zeros_like(bt)
            torch.cuda.synchronize(logr.device)
            it = model.solver.solve_device(bt.data_ptr(), p.data_ptr(), tol=model.tol)
        else:
            x, it = model.solver.solve(b, tol=model.tol)
            p = torch.from_numpy(x)
        if it <= 0:
            raise RuntimeError(f"pressure solve failed ({it})")
        ctx.model, ctx.g, ctx.val = model, g, val
        ctx.save_for_backward(p)
        return p

    @staticmethod
    def backward(ctx, grad_p):
        (p,) = ctx.saved_tensors
        model, net = ctx.model, ctx.model.net
        model.refresh(ctx.g, ctx.val)  # the matrix of this forward
        grad_p = grad_p.contiguous()
        if grad_p.is_cuda:
            dev = grad_p.device
            eu, ev, pdir = model.device_edges(dev)
            gt = torch.from_numpy(ctx.g).to(dev)
            lam = torch.zeros_like(p)
            out = torch.empty(net.n_edges, dtype=torch.float64, device=dev)
            torch.cuda.synchronize(dev)
            it = model.solver.adjoint_grad_device(p.data_ptr(), grad_p.data_ptr(), net.n_edges, eu.data_ptr(),
                                                  ev.data_ptr(), gt.data_ptr(), pdir.data_ptr(), net.n_dirichlet,
                                                  lam.data_ptr(), None, out.data_ptr(), tol=model.tol)
        else:
            _, _, dlr, it = model.solver.adjoint_grad(p.numpy(), grad_p.numpy(), net.eu, net.ev, ctx.g,
                                                      net.p_dirichlet, tol=model.tol)
            out = torch.from_numpy(dlr)
        if it <= 0:
            raise RuntimeError(f"adjoint solve failed ({it})")
        return out, None


def direct_gradient(net, logr, obs, p_obs):
    """FP64 direct-solve gradient of 1/2 sum (p[obs] - p_obs)^2 in log r."""
    g = net.conductance(np.exp(logr))
    val, b = net.assemble(g)
    A = sp.csr_matrix((val, net.ja, net.ia), shape=(net.n, net.n)).tocsc()
    lu = spla.splu(A)
    p = lu.solve(b)
    d = np.zeros(net.n)
    np.add.at(d, obs, p[obs] - p_obs)
    lam = lu.solve(d)
    pd = net.p_dirichlet
    pu = np.where(net.eu >= 0, p[np.maximum(net.eu, 0)], pd[np.maximum(-1 - net.eu, 0)])
    pv = np.where(net.ev >= 0, p[np.maximum(net.ev, 0)], pd[np.maximum(-1 - net.ev, 0)])
    lu_ = np.where(net.eu >= 0, lam[np.maximum(net.eu, 0)], 0.0)
    lv_ = np.where(net.ev >= 0, lam[np.maximum(net.ev, 0)], 0.0)
    return 0.5 * float((p[obs] - p_obs) @ (p[obs] - p_obs)), 4.0 * g * (-(lu_ - lv_) * (pu - pv))


def main():
    t0 = time.perf_counter()
    out = {"ok": False}
    net = make_network(12, 6, seed=1)
    rng = np.random.default_rng(11)
    logr0 = np.log(net.radius)
    logr_true = logr0 + rng.normal(0.0, 0.1, net.n_edges)
    obs = rng.choice(net.n, size=200, replace=False)
    gt = net.conductance(np.exp(logr_true))
    vt, bt = net.assemble(gt)
    p_obs = spla.spsolve(sp.csr_matrix((vt, net.ja, net.ia), shape=(net.n, net.n)).tocsc(), bt)[obs]
    say(f"[harness] network n={net.n} segments={net.n_edges}; torch {torch.__version__}")
    f_ref, g_ref = direct_gradient(net, logr0, obs, p_obs)
    model = NetworkModel(net, min_size=300)
    obs_t = torch.from_numpy(obs)

    def loss_of(logr_t):
        p = PressureSolve.apply(logr_t, model)
        r = p[obs_t.to(p.device)] - torch.from_numpy(p_obs).to(p.device)
        return 0.5 * (r * r).sum()

    # CPU tensors (host entry points)
    x = torch.tensor(logr0, dtype=torch.float64, requires_grad=True)
    f = loss_of(x)
    f.backward()
    g_cpu = x.grad.numpy().copy()
    err_ref = float(np.linalg.norm(g_cpu - g_ref) / np.linalg.norm(g_ref))
    # the C ABI's own gradient at the same point
    val, b = net.assemble(net.conductance(np.exp(logr0)))
    model.solver.update(net.ia, net.ja, val)
    model.current = None
    p, _ = model.solver.solve(b, tol=1e-13)
    d = np.zeros(net.n)
    np.add.at(d, obs, p[obs] - p_obs)
    _, _, g_abi, _ = model.solver.adjoint_grad(p, d, net.eu, net.ev, net.conductance(np.exp(logr0)),
                                               net.p_dirichlet, tol=1e-13)
    err_abi = float(np.abs(g_cpu - g_abi).max() / np.abs(g_abi).max())
    f = float(f.detach())
    out.update(loss=f, loss_direct=f_ref, cpu_vs_direct=err_ref, cpu_vs_abi=err_abi)
    say(f"[harness] CPU tensors: loss {f:.10e} (direct {f_ref:.10e}); gradient vs FP64 direct "
        f"{err_ref:.2e}, vs C ABI {err_abi:.2e}")
    ok = abs(f - f_ref) <= 1e-9 * f_ref and err_ref < 1e-6 and err_abi < 1e-12

    # CUDA tensors (device-pointer entry points)
    if torch.cuda.is_available():
        xc = torch.tensor(logr0, dtype=torch.float64, device="cuda", requires_grad=True)
        fc = loss_of(xc)
        fc.backward()
        g_cuda = xc.grad.cpu().numpy()
        err_dev = float(np.linalg.norm(g_cuda - g_cpu) / np.linalg.norm(g_cpu))
        fc = float(fc.detach())
        out.update(cuda_loss=fc, cuda_vs_cpu=err_dev)
        say(f"[harness] CUDA tensors: loss {fc:.10e}; gradient vs CPU tensors {err_dev:.2e}")
        ok = ok and abs(fc - f) <= 1e-10 * f and err_dev < 1e-9
    else:
        out["cuda"] = "torch.cuda not available"
        ok = False

    # an optimizer on the log-radii
    xo = torch.tensor(logr0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([xo], lr=0.01)
    hist = []
    for _ in range(10):
        opt.zero_grad()
        fo = loss_of(xo)
        fo.backward()
        hist.append(float(fo.detach()))
        opt.step()
    hist.append(float(loss_of(xo.detach())))
    out["adam_losses"] = hist
    say(f"[harness] Adam on log r: {hist[0]:.4e} -> {hist[-1]:.4e}")
    ok = ok and hist[-1] < 0.5 * hist[0]
    out["ok"] = bool(ok)
    out["seconds"] = round(time.perf_counter() - t0, 1)
    model.solver.free()
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
