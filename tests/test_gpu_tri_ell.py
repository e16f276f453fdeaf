"""The exact IC(0) triangular solves (C_amg1/src/ichol.c:121-165) from the
level-ordered ELL copies against the same solves from the CSR rows with one
lane per row: every row's sum is formed in the same order, so the
preconditioner, every PCG iterate and the solution must agree to the last
bit (FP64 and FP32, fine factor and the coarsest IC-PCG's factor), and the
ELL copies must follow the factors through a value-only rebuild on the
device.  Inside a PCG solve the ELL levels run as one launch per run of
levels (bn_trsv_walk, rows waiting on their predecessors' values); outside
it, and with BN_TRI_WALK=0, one launch per level.  Both must agree to the
last bit, and a walk that gives up waiting must hand the solve back to the
launch-per-level form with the same result."""
import sys

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def net():
    # above 2^17 rows, so the untuned CSR rows of the factor run one lane per row
    return make_network(52, 10, seed=2)


def solver(monkeypatch, net, ell, walk=True, **kw):
    monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
    monkeypatch.setenv("BN_TRI_ELL", "1" if ell else "0")
    monkeypatch.setenv("BN_TRI_WALK", "1" if walk else "0")
    return AMGSolver(net.ia, net.ja, net.val, 0.25, 500, 20, 1, coarsening="pmis", **kw)


MODES = {"fp64": dict(), "fp64_icpcg_coarsest": dict(coarse="icpcg"), "fp32_exact_ic": dict(precision="fp32")}


@pytest.mark.parametrize("mode", list(MODES))
def test_ell_solves_equal_the_csr_rows(monkeypatch, net, mode):
    case(f"exact IC(0) from ELL copies vs CSR rows, {mode}, n={net.n}")
    out = []
    for ell in (True, False):
        s = solver(monkeypatch, net, ell, **MODES[mode])
        rng = np.random.default_rng(5)
        g = rng.standard_normal(net.n)
        z = s.precondition(g, "ichol")
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)
        out.append((z, x, it))
        s.free()
    (z1, x1, i1), (z0, x0, i0) = out
    print(f"iterations {i1} / {i0}; max |dz| {np.max(np.abs(z1 - z0)):.3e}", flush=True)
    assert i1 > 0 and i1 == i0
    assert np.array_equal(z1, z0)
    assert np.array_equal(x1, x0)


def test_ell_copies_follow_a_device_rebuild(monkeypatch, net):
    case(f"ELL copies after bn_solver_rebuild_values_device, n={net.n}")
    from blood_network_amd import _lib as L
    val2 = net.val * (1.0 + 0.1 * np.random.default_rng(3).random(net.nnz))
    # symmetric scaling keeps the matrix symmetric: scale each edge once
    A = __import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix((val2, net.ja, net.ia), shape=(net.n, net.n))
    A = ((A + A.T) * 0.5).tocsr()
    A.sort_indices()
    assert np.array_equal(A.indptr, net.ia) and np.array_equal(A.indices, net.ja)
    res = []
    for ell in (True, False):
        s = solver(monkeypatch, net, ell)
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)   # builds the ELL copies
        d = L.DeviceArray(A.data.copy())
        s.rebuild_values_device(net.nnz, d.ptr)
        x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)
        res.append((x2, it2))
        d.free()
        s.free()
    (xa, ia_), (xb, ib) = res
    print(f"after the rebuild: iterations {ia_} / {ib}", flush=True)
    assert ia_ == ib and np.array_equal(xa, xb)


@pytest.mark.parametrize("mode", list(MODES))
def test_walk_equals_a_launch_per_level(monkeypatch, capfd, net, mode):
    case(f"exact IC(0): one launch per run of levels vs one per level, {mode}, n={net.n}")
    out = []
    for walk in (True, False):
        s = solver(monkeypatch, net, True, walk=walk, **MODES[mode])
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)
        x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)   # replayed step graphs
        out.append((x, it, x2, it2))
        s.free()
    (xw, iw, xw2, iw2), (xl, il, xl2, il2) = out
    print(f"iterations walk {iw}/{iw2}, per level {il}/{il2}", flush=True)
    assert "timed out" not in capfd.readouterr().err  # the walk itself ran, no hand-back
    assert iw > 0 and iw == il == iw2 == il2
    assert np.array_equal(xw, xl) and np.array_equal(xw2, xl2) and np.array_equal(xw, xw2)


def test_walk_that_gives_up_hands_back_to_a_launch_per_level(monkeypatch, capfd, net):
    case(f"exact IC(0) walk forced to give up at its first wait (BN_TRI_WALK_LIMIT=-1), n={net.n}")
    ref = solver(monkeypatch, net, True, walk=False)
    x0, i0 = ref.solve(net.b, tol=1e-10, maxiter=500)
    ref.free()
    monkeypatch.setenv("BN_TRI_WALK_LIMIT", "-1")
    s = solver(monkeypatch, net, True, walk=True)
    x1, i1 = s.solve(net.b, tol=1e-10, maxiter=500)
    x2, i2 = s.solve(net.b, tol=1e-10, maxiter=500)   # the solver keeps the launch per level
    s.free()
    err = capfd.readouterr().err
    print(f"iterations {i1}/{i2} vs {i0}", flush=True)
    assert "timed out" in err
    assert i1 == i2 == i0 and np.array_equal(x1, x0) and np.array_equal(x2, x0)
