"""The PCG step's p'Ap fused into the row-block product q = A p against the
product followed by its own dot pass (BN_FUSE_DOT=0).  Only the order of
the dot's sums differs (a partial per row block instead of per strided
slot), so both solves take the same iterations
(C_amg1/src/amg_pcg.c:56-59: alpha = r'z / p'Ap) and reach the same
solution to rounding; the fused solve is deterministic from one graph
replay to the next."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import AMGSolver, make_network

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_pap_equals_separate_dot(monkeypatch, precision):
    net = make_network(52, 10, seed=2)
    case(f"fused p'Ap vs separate dot, {precision}, n={net.n}")
    kw = dict(coarsening="hmis", precision=precision)
    if precision == "fp32":
        kw.update(fine_ic="jacobi", ic_sweeps=4, max_colours=16, stop="relres")
    out = []
    for fuse in ("1", "0"):
        # every operator on the row-block kernel, so the fused path runs
        monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
        monkeypatch.setenv("BN_APPLY", "blocks")
        monkeypatch.setenv("BN_FUSE_DOT", fuse)
        s = AMGSolver(net.ia, net.ja, net.val, 0.25, 500, 20, 1, coarse="dense", **kw)
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)
        x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)   # replayed step graphs
        out.append((x, it, x2, it2))
        s.free()
    (xf, itf, xf2, itf2), (xs, its, xs2, its2) = out
    print(f"iterations fused {itf}/{itf2}, separate {its}/{its2}; "
          f"max rel diff {np.abs(xf - xs).max() / np.abs(xs).max():.2e}", flush=True)
    assert itf > 0 and itf == its and itf2 == itf and its2 == its
    assert np.array_equal(xf.view(np.uint64), xf2.view(np.uint64))
    assert np.abs(xf - xs).max() <= 1e-8 * np.abs(xs).max()
    # the two orders of the dot's sums round differently, so equal bits
    # would mean that the fused product never ran
    assert not np.array_equal(xf, xs)
    # the solution itself: residual of the fused solve
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    A = 0.5 * (A + A.T)
    assert np.linalg.norm(net.b - A @ xf) <= 1e-7 * np.linalg.norm(net.b)
