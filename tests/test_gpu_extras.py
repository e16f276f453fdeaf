"""North-star forms beside the plain solve: the adjoint gradient through
the same hierarchy, and batched right-hand sides."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from blood_network_amd import AMGSolver, make_network

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def direct_solve(net, g):
    val, b = net.assemble(g)
    A = sp.csr_matrix((val, net.ja, net.ia), shape=(net.n, net.n)).tocsc()
    return A, spla.spsolve(A, b)


def loss(net, logr, obs, p_obs):
    g = net.conductance(np.exp(logr))
    _, p = direct_solve(net, g)
    return 0.5 * float(np.sum((p[obs] - p_obs) ** 2))


@pytest.fixture(scope="module")
def adjoint_case():
    net = make_network(12, 6, seed=1)
    rng = np.random.default_rng(11)
    logr0 = np.log(net.radius)
    logr_true = logr0 + rng.normal(0.0, 0.1, net.n_edges)
    obs = rng.choice(net.n, size=200, replace=False)
    _, p_true = direct_solve(net, net.conductance(np.exp(logr_true)))
    return net, logr0, obs, p_true[obs]


def test_adjoint_gradient_matches_fd_and_cpu_reference(adjoint_case):
    case("adjoint gradient vs central FD and FP64 CPU gradient")
    net, logr, obs, p_obs = adjoint_case
    g = net.conductance(np.exp(logr))
    val, b = net.assemble(g)
    s = AMGSolver(net.ia, net.ja, val, min_size=300)
    p, it = s.solve(b, tol=1e-13, maxiter=500)
    dLdp = np.zeros(net.n)
    dLdp[obs] = p[obs] - p_obs
    lam, dLdg, dLdlogr, it_adj = s.adjoint_grad(p, dLdp, net.eu, net.ev, g, net.p_dirichlet, tol=1e-13)
    assert it > 0 and it_adj > 0
    # FP64 CPU reference: direct solves, same formula
    A, p_ref = direct_solve(net, g)
    d = np.zeros(net.n)
    d[obs] = p_ref[obs] - p_obs
    lam_ref = spla.spsolve(A, d)
    pu = np.where(net.eu >= 0, p_ref[np.maximum(net.eu, 0)], net.p_dirichlet[np.maximum(-1 - net.eu, 0)])
    pv = np.where(net.ev >= 0, p_ref[np.maximum(net.ev, 0)], net.p_dirichlet[np.maximum(-1 - net.ev, 0)])
    lu = np.where(net.eu >= 0, lam_ref[np.maximum(net.eu, 0)], 0.0)
    lv = np.where(net.ev >= 0, lam_ref[np.maximum(net.ev, 0)], 0.0)
    ref = 4.0 * g * (-(lu - lv) * (pu - pv))
    err = np.linalg.norm(dLdlogr - ref) / np.linalg.norm(ref)
    print(f"adjoint its={it_adj} rel err vs FP64 CPU gradient={err:.2e}")
    assert err < 1e-6
    # central finite differences in log r, relative step 1e-6
    rng = np.random.default_rng(3)
    big = np.argsort(-np.abs(ref))[:2000]
    edges = rng.choice(big, size=16, replace=False)
    h = 1e-6
    for e in edges:
        lp, lm = logr.copy(), logr.copy()
        lp[e] += h
        lm[e] -= h
        fd = (loss(net, lp, obs, p_obs) - loss(net, lm, obs, p_obs)) / (2 * h)
        assert abs(fd - dLdlogr[e]) <= 1e-4 * abs(fd) + 1e-6 * np.abs(ref).max(), (e, fd, dLdlogr[e])


def test_batch_equals_single_solves():
    case("batched right-hand sides")
    net = make_network(16, 9)
    s = AMGSolver(net.ia, net.ja, net.val)
    rng = np.random.default_rng(2)
    B = np.stack([net.rhs_for(np.array([rng.uniform(8e3, 16e3), rng.uniform(0.5e3, 2e3)])) for _ in range(4)])
    X, its = s.solve_batch(B, tol=1e-10)
    for j in range(4):
        xj, ij = s.solve(B[j], tol=1e-10)
        assert its[j] == ij and np.array_equal(X[j], xj)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    res = np.linalg.norm(B.T - A @ X.T, axis=0) / np.linalg.norm(B, axis=1)
    assert res.max() < 1e-9
