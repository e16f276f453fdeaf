"""Block PCG (north-star batching form): k right-hand sides in one solve
over CSR SpMM with interleaved columns.  Kernels against the oracle per
column; every block column against its single-right-hand-side solve."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import AMGSolver, make_network
from blood_network_amd import _lib as L
from oracle_lib import dp, ip
from test_gpu_kernels import coarse_level_matrix, dev, hub_matrix, relerr, seg_bin

pytestmark = pytest.mark.gpu
lib = L.lib
lib.bn_spmm.argtypes = [C.c_char, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                        C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                        C.c_int, C.c_void_p, C.c_void_p]
lib.bn_gs_color_map_k.argtypes = [C.c_char, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("vt", ["d", "s"])
@pytest.mark.parametrize("K", [1, 4, 32])
def test_spmm_against_oracle(oracle, vt, K):
    s, net = coarse_level_matrix()
    mats = [("A%d" % lev, s.level(lev)["A"]) for lev in range(s.num_levels())]
    mats.append(("hubs", hub_matrix(2)))
    rng = np.random.default_rng(10)
    dt = np.float64 if vt == "d" else np.float32
    tol = 1e-12 if vt == "d" else 1e-5
    for name, (ia, ja, val) in mats:
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        n = len(ia) - 1
        rows, seg_ptr, seg_row, nmulti = seg_bin(ia, 64)
        case(f"spmm {name} n={n} K={K} long={len(rows)} multi={nmulti} vt={vt}")
        X, Bm, Y0 = rng.random((n, K)), rng.random((n, K)), rng.random((n, K))
        one = np.zeros(1, np.int32)
        dia, dja, dv = dev(ia), dev(ja), dev(val.astype(dt))
        dX, dB = dev(X.astype(dt)), dev(Bm.astype(dt))
        drows, dsp, dsr = dev(rows if len(rows) else one), dev(seg_ptr), dev(seg_row if len(seg_row) else one)
        dpart = dev(np.zeros(max(1, int(seg_ptr[nmulti]) * K)))
        for op in (0, 1, 2):
            dY = dev(Y0.astype(dt) if op == 2 else np.zeros((n, K), dt))
            lib.bn_spmm(vt.encode(), K, op, n, dia.ptr, dja.ptr, dv.ptr, dX.ptr, dY.ptr, dB.ptr, 64, len(rows),
                        drows.ptr, dsp.ptr, dsr.ptr, int(seg_ptr[-1]), nmulti, dpart.ptr, None)
            got = dY.numpy().astype(np.float64).reshape(n, K)
            for c in range(K):
                ref = np.zeros(n)
                if op == 0:
                    oracle.lib.abyvg(ip(ia), ip(ja), dp(val), dp(np.ascontiguousarray(X[:, c])), n, dp(ref))
                elif op == 1:
                    oracle.lib.abyvam(dp(np.ascontiguousarray(Bm[:, c])), ip(ia), ip(ja), dp(val),
                                      dp(np.ascontiguousarray(X[:, c])), n, dp(ref))
                else:
                    ref = np.ascontiguousarray(Y0[:, c]).copy()
                    oracle.lib.abyvcs(ip(ia), ip(ja), dp(val), dp(np.ascontiguousarray(X[:, c])), n, dp(ref))
                assert relerr(got[:, c], ref) < tol, (name, op, c)


@pytest.mark.parametrize("vt", ["d", "s"])
def test_block_gs_colours_are_exact_gauss_seidel(vt):
    # the K-column colour kernels: every column satisfies the closed form
    # (D + L_pi) x_new = b - U_pi x_old of a sweep in colour order
    net = make_network(14, 8, seed=3)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    n, K = net.n, 8
    m = L.CSR(n, n, net.nnz, ip(net.ia), ip(net.ja), dp(net.val))
    perm = np.zeros(n, np.int32)
    cptr = C.POINTER(C.c_int)()
    mis = (np.arange(n) % 3 == 0).astype(np.int32)
    ncol = lib.bn_colour_order(C.byref(m), ip(mis), 1, ip(perm), C.byref(cptr))
    col_ptr = np.ctypeslib.as_array(cptr, (ncol + 1,)).copy()
    Ac = A[perm].tocsr()
    Ac.sort_indices()
    pos = np.empty(n, np.int64)
    pos[perm] = np.arange(n)
    rng = np.random.default_rng(5)
    Bm = rng.random((n, K)) * 1e-9
    X0 = rng.random((n, K)) * 1e3
    D = A.diagonal()
    dt = np.float64 if vt == "d" else np.float32
    case(f"block GS colours n={n} K={K} colours={ncol} vt={vt}")
    dia, dja, dv = dev(Ac.indptr.astype(np.int32)), dev(Ac.indices.astype(np.int32)), dev(Ac.data.astype(dt))
    dmap, ddinv = dev(perm), dev((1.0 / D[perm]).astype(dt))
    dB, dX = dev(Bm.astype(dt)), dev(X0.astype(dt))
    for c in range(ncol):
        lib.bn_gs_color_map_k(vt.encode(), K, int(col_ptr[c]), int(col_ptr[c + 1]), dmap.ptr, dia.ptr, dja.ptr,
                              dv.ptr, ddinv.ptr, dB.ptr, dX.ptr, None, None, None)
    Xn = dX.numpy().astype(np.float64).reshape(n, K)
    coo = A.tocoo()
    lo = pos[coo.col] < pos[coo.row]
    up = pos[coo.col] > pos[coo.row]
    Lp = sp.csr_matrix((coo.data[lo], (coo.row[lo], coo.col[lo])), shape=(n, n))
    Up = sp.csr_matrix((coo.data[up], (coo.row[up], coo.col[up])), shape=(n, n))
    for c in range(K):
        lhs = D * Xn[:, c] + Lp @ Xn[:, c]
        rhs = Bm[:, c] - Up @ X0[:, c]
        assert relerr(lhs, rhs) < (1e-12 if vt == "d" else 2e-5)


def rhs_batch(net, k, seed):
    rng = np.random.default_rng(seed)
    return np.stack([net.rhs_for(np.array([rng.uniform(8e3, 16e3), rng.uniform(0.5e3, 2e3)])) for _ in range(k)])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("k", [3, 32])
def test_block_columns_equal_single_solves(precision, k):
    net = make_network(16, 9)
    case(f"block PCG lattice16 k={k} {precision}")
    s = AMGSolver(net.ia, net.ja, net.val, precision=precision, fine_ic="jacobi", flexible=True, min_size=400)
    B = rhs_batch(net, k, 7)
    B[1] = 0.0                                   # a zero right-hand side stops at once
    X, its = s.solve_block(B, tol=1e-10)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    assert its[1] == 0 and not np.any(X[1])
    for j in range(k):
        if j == 1:
            continue
        xj, ij = s.solve(B[j], tol=1e-10)
        assert abs(int(its[j]) - ij) <= 1, (j, its[j], ij)
        assert relerr(X[j], xj) < 1e-8, j
        assert np.linalg.norm(B[j] - A @ X[j]) / np.linalg.norm(B[j]) < 1e-8
    print(f"iterations {its.tolist()}, block solve {s.report.solve_ms:.2f} ms")


def test_block_deep_hierarchy_and_graph_replay(net30, monkeypatch):
    # HMIS hierarchy with hub rows on the coarse levels (segmented bin);
    # graph replays are bit-identical to eager launches
    case("block PCG lattice30 hmis k=8, graph vs eager")
    s = AMGSolver(net30.ia, net30.ja, net30.val, coarsening="hmis", min_size=300, precision="fp32",
                  fine_ic="jacobi")
    B = rhs_batch(net30, 8, 9)
    monkeypatch.setenv("BN_NO_GRAPH", "1")
    Xe, ie = s.solve_block(B, tol=1e-9)
    monkeypatch.delenv("BN_NO_GRAPH")
    for _ in range(2):
        Xg, ig = s.solve_block(B, tol=1e-9)
        assert np.array_equal(ig, ie) and np.array_equal(Xg, Xe)
    A = sp.csr_matrix((net30.val, net30.ja, net30.ia), shape=(net30.n, net30.n))
    res = np.linalg.norm(B.T - A @ Xg.T, axis=0) / np.linalg.norm(B, axis=1)
    assert res.max() < 1e-8
    x0, i0 = s.solve(B[0], tol=1e-9)
    assert relerr(Xg[0], x0) < 1e-7


def test_block_rejects_unsupported_modes():
    net = make_network(8, 5)
    s = AMGSolver(net.ia, net.ja, net.val, min_size=100)      # exact fine IC
    with pytest.raises(RuntimeError, match="IC"):
        s.solve_block(np.stack([net.b, net.b]))
    s2 = AMGSolver(net.ia, net.ja, net.val, min_size=100, fine_ic="jacobi")
    with pytest.raises(RuntimeError):
        s2.solve_block(np.zeros((33, net.n)))
