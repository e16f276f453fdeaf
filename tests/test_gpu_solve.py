"""End-to-end parity of the GPU solve with the oracle (the reference's
algorithm restated on the CPU) and with an FP64 sparse direct solve.

Bar (from the solve's contract): converged pressures within 1e-8 relative
2-norm of the reference solution and of the direct solve; PCG iteration
counts reported beside the reference's (the multicolour smoother is a
different Gauss-Seidel ordering, so counts may differ by a few).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from blood_network_amd import AMGSolver, make_network

pytestmark = pytest.mark.gpu


def direct(net):
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    return spla.spsolve(A.tocsc(), net.b)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("coarse", ["dense", "icpcg"])
def test_lattice16_rs_matches_reference(net16, oracle, coarse):
    _case(f"lattice16 RS coarse={coarse}")
    xo, it_o, _ = oracle.build(net16.ia, net16.ja, net16.val, coarsening="rs").solve(net16.b, 1e-10, 1000)
    s = AMGSolver(net16.ia, net16.ja, net16.val, coarsening="rs", coarse=coarse)
    x, it = s.solve(net16.b, tol=1e-10, maxiter=1000)
    xd = direct(net16)
    print(f"iterations gpu={it} reference={it_o} rel(x, ref)={rel(x, xo):.2e} rel(x, direct)={rel(x, xd):.2e}")
    assert it > 0
    assert abs(it - it_o) <= 3
    assert rel(x, xo) < 1e-8
    assert rel(x, xd) < 1e-8


@pytest.mark.parametrize("coarsening", ["pmis", "hmis"])
def test_lattice30_matches_reference(net30, oracle, coarsening):
    _case(f"lattice30 {coarsening}")
    xo, it_o, _ = oracle.build(net30.ia, net30.ja, net30.val, coarsening=coarsening, min_size=500).solve(
        net30.b, 1e-10, 1000)
    s = AMGSolver(net30.ia, net30.ja, net30.val, coarsening=coarsening, min_size=500)
    assert s.num_levels() >= 3
    x, it = s.solve(net30.b, tol=1e-10, maxiter=1000)
    xd = direct(net30)
    print(f"iterations gpu={it} reference={it_o} rel(x, ref)={rel(x, xo):.2e} rel(x, direct)={rel(x, xd):.2e}")
    assert abs(it - it_o) <= 4
    assert rel(x, xo) < 1e-8
    assert rel(x, xd) < 1e-8


def test_fp32_vcycle_flexible(net30):
    _case("lattice30 FP32 V-cycle, flexible PCG")
    xd = direct(net30)
    s64 = AMGSolver(net30.ia, net30.ja, net30.val, coarsening="hmis", min_size=500, stop="relres")
    x64, it64 = s64.solve(net30.b, tol=1e-12, maxiter=1000)
    s32 = AMGSolver(net30.ia, net30.ja, net30.val, coarsening="hmis", min_size=500, precision="fp32",
                    stop="relres")
    x32, it32 = s32.solve(net30.b, tol=1e-8, maxiter=1000)
    A = sp.csr_matrix((net30.val, net30.ja, net30.ia), shape=(net30.n, net30.n))
    relres = np.linalg.norm(net30.b - A @ x32) / np.linalg.norm(net30.b)
    print(f"fp32 iters={it32} fp64 iters={it64} relres={relres:.2e} err vs fp64={rel(x32, x64):.2e}")
    assert it32 > 0
    assert relres < 1e-8 * 1.01
    # the error at a 1e-8 residual is ~cond * 1e-8: reported, loosely bounded
    assert rel(x64, xd) < 1e-8
    assert rel(x32, x64) < 1e-3


def test_jacobi_ic_flexible(net30):
    _case("lattice30 Jacobi IC, flexible")
    s = AMGSolver(net30.ia, net30.ja, net30.val, coarsening="pmis", min_size=500, fine_ic="jacobi",
                  stop="relres")
    # the error bound is kappa times the residual bound: solved to 1e-12 so
    # that the 1e-8 error test holds for kappa up to 1e4
    x, it = s.solve(net30.b, tol=1e-12, maxiter=1000)
    assert it > 0
    assert rel(x, direct(net30)) < 1e-8


def test_initial_guess_and_zero_rhs(net16):
    _case("lattice16 initial guess / zero rhs")
    s = AMGSolver(net16.ia, net16.ja, net16.val)
    xd = direct(net16)
    x0, _ = s.solve(net16.b, tol=1e-2, maxiter=100)
    x, it = s.solve(net16.b, x=x0, tol=1e-6, maxiter=100)
    # x is used as the initial guess (amg_wrapper.c:207-208): the second
    # solve continues from x0 and ends closer to the solution
    assert it > 0 and rel(x, xd) < 0.1 * rel(x0, xd)
    z, it0 = s.solve(np.zeros(net16.n), tol=1e-10, maxiter=10)
    assert it0 == 0 and np.all(z == 0)


def test_maxiter_failure_returns_negative(net16):
    _case("lattice16 maxiter=1")
    s = AMGSolver(net16.ia, net16.ja, net16.val)
    _, it = s.solve(net16.b, tol=1e-14, maxiter=1)
    assert it == -1


@pytest.mark.parametrize("apply", ["lanes", "blocks"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_operator_kernel_choice_matches_reference(net30, oracle, apply, precision, monkeypatch):
    # level operators pinned to the lanes-per-row or the row-block kernel
    # (normally chosen by timing at upload): both give the reference solve
    _case(f"lattice30 hmis apply={apply} {precision}")
    monkeypatch.setenv("BN_APPLY", apply)
    xo, it_o, _ = oracle.build(net30.ia, net30.ja, net30.val, coarsening="hmis", min_size=300).solve(
        net30.b, 1e-10, 1000)
    s = AMGSolver(net30.ia, net30.ja, net30.val, coarsening="hmis", min_size=300, precision=precision,
                  fine_ic="jacobi" if precision == "fp32" else "exact")
    x, it = s.solve(net30.b, tol=1e-10, maxiter=1000)
    monkeypatch.delenv("BN_APPLY")
    print(f"iterations gpu={it} reference={it_o} rel(x, ref)={rel(x, xo):.2e}")
    assert it > 0
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("coarsening", ["pmis", "hmis"])
def test_block_smoother_matches_reference(net30, oracle, coarsening, precision):
    # chunked l1 block Gauss-Seidel on every level instead of the
    # multicolour sweep: a different smoother, the same converged solve
    _case(f"lattice30 {coarsening} block smoother {precision}")
    xo, it_o, _ = oracle.build(net30.ia, net30.ja, net30.val, coarsening=coarsening, min_size=300).solve(
        net30.b, 1e-10, 1000)
    s = AMGSolver(net30.ia, net30.ja, net30.val, coarsening=coarsening, min_size=300, precision=precision,
                  fine_ic="jacobi" if precision == "fp32" else "exact", block_smoother=True, chunk_rows=512)
    x, it = s.solve(net30.b, tol=1e-10, maxiter=1000)
    print(f"iterations gpu={it} reference={it_o} rel(x, ref)={rel(x, xo):.2e}")
    assert 0 < it <= 2 * it_o + 5
    assert rel(x, xo) < 1e-8


def test_solve_is_deterministic(net30):
    _case("lattice30 bitwise reproducibility")
    s = AMGSolver(net30.ia, net30.ja, net30.val, min_size=500)
    x1, i1 = s.solve(net30.b, tol=1e-10)
    x2, i2 = s.solve(net30.b, tol=1e-10)
    assert i1 == i2 and np.array_equal(x1, x2)


@pytest.mark.parametrize("block", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replay_equals_eager(net30, precision, block, monkeypatch):
    # every PCG step replays one captured hipGraph; eager launches of the
    # same kernels (BN_NO_GRAPH=1, same solver, so the same autotuned
    # launch shapes) must give bit-identical iterates, across several
    # solves with other work in between
    _case(f"lattice30 graph vs eager {precision} block={block}")
    kw = dict(coarsening="pmis", min_size=500, precision=precision,
              fine_ic="jacobi" if precision == "fp32" else "exact", block_smoother=block, chunk_rows=512)
    s = AMGSolver(net30.ia, net30.ja, net30.val, **kw)
    monkeypatch.setenv("BN_NO_GRAPH", "1")
    xe, ie = s.solve(net30.b, tol=1e-10)
    monkeypatch.delenv("BN_NO_GRAPH")
    for rep in range(3):
        xg, ig = s.solve(net30.b, tol=1e-10)
        assert ig == ie and np.array_equal(xg, xe)
        monkeypatch.setenv("BN_NO_GRAPH", "1")
        s.solve(net30.b * (rep + 2.0), tol=1e-6)  # eager work between replays
        monkeypatch.delenv("BN_NO_GRAPH")
