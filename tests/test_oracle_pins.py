"""Pin the oracle before trusting it.

1. Against the compiled reference (oracle/_ref, built from the C_amg1
   sources): every hierarchy array and the converged x bit for bit.
2. Against mathematics that fixes each routine independently of the
   reference's code: closed forms, invariants, scipy on small cases.
   Each check is chosen so a plausible slip (dropped term, wrong sign,
   transposed operand, wrong order) fails it.

The reference ships no golden vectors for this path (its only published
numbers use matrix files it does not ship), so these are the pins.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from blood_network_amd import make_network
from oracle_lib import dp, ip


def csr(t, shape=None):
    ia, ja, v = t
    n = len(ia) - 1
    return sp.csr_matrix((v, ja, ia), shape=shape or (n, int(ja.max()) + 1 if len(ja) else 0))


def random_graph_laplacian(n, deg, seed, dirichlet=3):
    """SPD weighted graph Laplacian with a few grounded nodes; unsorted
    columns in some rows to exercise order dependence."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        nb = rng.choice(n, size=deg, replace=False)
        for j in nb:
            if j == i:
                continue
            w = float(np.exp(rng.normal(0, 2)))
            rows += [i, j]
            cols += [j, i]
            vals += [-w, -w]
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    d = -np.asarray(A.sum(axis=1)).ravel()
    d[rng.choice(n, size=dirichlet, replace=False)] += 1.0
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    return A


def hier_equal(a, b, nl):
    for lev in range(nl):
        x, y = a.level(lev), b.level(lev)
        assert x.keys() == y.keys()
        for k in x:
            if isinstance(x[k], tuple):
                for p, q in zip(x[k], y[k]):
                    assert np.array_equal(p, q), (lev, k)
            else:
                assert np.array_equal(x[k], y[k]), (lev, k)


# ---------------------------------------------------------------- 1. _ref
@pytest.mark.parametrize("coarsening,L,d,min_size", [("rs", 16, 9, 4096), ("pmis", 20, 8, 300),
                                                     ("hmis", 20, 8, 300), ("rs", 20, 8, 300)])
def test_oracle_equals_compiled_reference_on_networks(oracle, reference, coarsening, L, d, min_size):
    net = make_network(L, d, seed=3)
    ho = oracle.build(net.ia, net.ja, net.val, coarsening=coarsening, min_size=min_size)
    hr = reference.build(net.ia, net.ja, net.val, coarsening=coarsening, min_size=min_size)
    assert ho.num_levels() == hr.num_levels() >= 2
    hier_equal(ho, hr, ho.num_levels())
    for a, b in zip(ho.fine_ichol(), hr.fine_ichol()):
        assert np.array_equal(a, b)
    xo, io, _ = ho.solve(net.b, 1e-10, 300)
    xr, ir, _ = hr.solve(net.b, 1e-10, 300)
    assert io == ir and np.array_equal(xo, xr)


@pytest.mark.parametrize("smoother", [0, 1, 2, 3])
def test_oracle_equals_reference_all_smoothers(oracle, reference, smoother):
    A = random_graph_laplacian(400, 5, seed=11)
    kw = dict(coarsening="pmis", min_size=60, smoother=smoother)
    ho = oracle.build(A.indptr, A.indices, A.data, **kw)
    hr = reference.build(A.indptr, A.indices, A.data, **kw)
    b = np.random.default_rng(0).random(400)
    zo, zr = ho.precond(b), hr.precond(b)
    assert np.array_equal(zo, zr)
    assert np.array_equal(ho.vcycle(b), hr.vcycle(b))
    xo, io, _ = ho.solve(b, 1e-9, 200)
    xr, ir, _ = hr.solve(b, 1e-9, 200)
    assert io == ir and np.array_equal(xo, xr)


def test_oracle_equals_reference_value_rebuild(oracle, reference):
    net = make_network(12, 7, seed=5)
    ho = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=200)
    hr = reference.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=200)
    g2 = net.g * np.exp(np.random.default_rng(1).normal(0, 0.3, net.n_edges))
    v2, _ = net.assemble(g2)
    ho.update(v2)
    hr.update(v2)
    hier_equal(ho, hr, ho.num_levels())


def test_oracle_equals_reference_unsorted_and_nonsymmetric(oracle, reference):
    A = random_graph_laplacian(300, 4, seed=2).tolil()
    A[5, 7] = A[5, 7] - 0.3  # non-symmetric entry: the symmetrisation matters
    A = A.tocsr()
    ia, ja, v = A.indptr.copy(), A.indices.copy(), A.data.copy()
    for i in range(0, 300, 7):  # reverse some rows: row order matters for IC(0)
        s, e = ia[i], ia[i + 1]
        ja[s:e] = ja[s:e][::-1].copy()
        v[s:e] = v[s:e][::-1].copy()
    for c in ("rs", "pmis", "hmis"):
        ho = oracle.build(ia, ja, v, coarsening=c, min_size=40)
        hr = reference.build(ia, ja, v, coarsening=c, min_size=40)
        hier_equal(ho, hr, ho.num_levels())


# ---------------------------------------------------------------- 2. math
def test_strength_closed_form(oracle):
    A = random_graph_laplacian(200, 6, seed=4)
    h = oracle.build(A.indptr, A.indices, A.data, coarsening="pmis", min_size=50, symmetrize=False)
    S = csr(h.level(0)["S"], (200, 200))
    for i in range(200):
        row = A.getrow(i)
        off = [(j, a) for j, a in zip(row.indices, row.data) if j != i]
        mx = max(0.0, max(-a for _, a in off))
        strong = sorted(j for j, a in off if -a >= 0.25 * mx)
        assert sorted(S.getrow(i).indices.tolist()) == strong


def test_split_invariants(oracle):
    A = random_graph_laplacian(500, 5, seed=8)
    for c in ("rs", "pmis", "hmis"):
        h = oracle.build(A.indptr, A.indices, A.data, coarsening=c, min_size=100, symmetrize=False)
        lv = h.level(0)
        S = csr(lv["S"], (500, 500))
        cf = lv["mis"]
        # every F point has a strong C neighbour after the second pass
        for i in np.where(cf == 0)[0]:
            assert cf[S.getrow(i).indices].any()
        # cf_order: C ascending then F ascending
        co = lv["cf_order"]
        nc = int(cf.sum())
        assert np.array_equal(co[:nc], np.sort(np.where(cf == 1)[0]))
        assert np.array_equal(co[nc:], np.sort(np.where(cf == 0)[0]))


def test_interpolation_properties(oracle):
    # scaled by 1e12: the reference's absolute 1e-15 guards
    # (rs_classical.c:778,791) must not bite for the closed forms below
    net = make_network(12, 6, seed=1)
    h = oracle.build(net.ia, net.ja, net.val * 1e12, coarsening="pmis", min_size=300)
    lv = h.level(0)
    A = csr(lv["A"])
    P = csr(lv["P"], (net.n, int(lv["mis"].sum())))
    cf = lv["mis"]
    rows_sum = np.asarray(P.sum(axis=1)).ravel()
    # C rows inject exactly one 1
    assert np.all(P[cf == 1].data == 1.0)
    assert np.all(np.diff(P.indptr)[cf == 1] == 1)
    # M-matrix: non-negative weights; zero-row-sum rows interpolate constants
    assert P.data.min() >= -1e-12  # (rounding only; a sign slip gives O(1) negatives)
    zero_rs = np.abs(np.asarray(A.sum(axis=1)).ravel()) < 1e-12 * np.abs(A.diagonal())
    f = (cf == 0) & zero_rs
    assert f.sum() > 10
    assert np.allclose(rows_sum[f], 1.0, rtol=1e-12)
    # pattern of F rows = strong C neighbours (in S-row order)
    S = csr(lv["S"], (net.n, net.n))
    cidx = np.cumsum(cf) - 1
    for i in np.where(cf == 0)[0][:200]:
        sc = [j for j in S.getrow(i).indices if cf[j]]
        assert P.getrow(i).indices.tolist() == [cidx[j] for j in sc]


def test_interpolation_absolute_guard_quirk(oracle):
    # In SI units capillary rows have |a^_ii| ~ 1e-15: rows whose |a^_ii|
    # <= 1e-15 keep unscaled numerators (rs_classical.c:791-794), so their
    # weights are O(a_ij) ~ 1e-16 instead of O(1).
    net = make_network(12, 6, seed=1)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=300)
    lv = h.level(0)
    cf = lv["mis"]
    P = csr(lv["P"], (net.n, int(cf.sum())))
    rs = np.asarray(P.sum(axis=1)).ravel()[cf == 0]
    tiny = np.abs(rs) < 1e-12
    assert tiny.sum() > 0 and (np.abs(rs[~tiny] - 1) < 0.5).mean() > 0.5


def test_galerkin_product(oracle):
    net = make_network(12, 6, seed=1)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="rs", min_size=300)
    l0, l1 = h.level(0), h.level(1)
    A = csr(l0["A"])
    nc = int(l0["mis"].sum())
    P = csr(l0["P"], (net.n, nc))
    Ac = csr(l1["A"], (nc, nc)).toarray()
    ref = (P.T @ A @ P).toarray()
    assert np.allclose(Ac, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.array_equal(Ac, Ac.T)
    # every coarse row starts with its diagonal (SMMP layout)
    ia, ja, _ = l1["A"]
    assert np.all(ja[ia[:-1]] == np.arange(nc))


def test_ichol_invariants(oracle):
    # sorted rows: (L L^T)_ij == a_ij on the pattern of tril(A)
    A = random_graph_laplacian(150, 4, seed=6)
    h = oracle.build(A.indptr, A.indices, A.data, coarsening="pmis", min_size=1000, symmetrize=False)
    L = csr(h.fine_ichol(), (150, 150))
    LLt = (L @ L.T).toarray()
    T = sp.tril(A).tocoo()
    assert np.allclose(LLt[T.row, T.col], T.data, rtol=1e-12, atol=1e-14)
    # diagonal-first rows (coarse levels): L = (D + L_A) D^{-1/2}
    net = make_network(12, 6, seed=2)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=800)
    last = h.level(h.num_levels() - 1)
    Ac = csr(last["A"])
    Lc = csr(last["Lc"], Ac.shape).toarray()
    D = Ac.diagonal()
    expect = np.tril(Ac.toarray(), -1) / np.sqrt(D)[None, :] + np.diag(np.sqrt(D))
    assert np.allclose(Lc, expect, rtol=1e-14, atol=0)


def test_ichol_solve_and_gs_against_scipy(oracle):
    lib = oracle.lib
    A = random_graph_laplacian(120, 4, seed=9)
    h = oracle.build(A.indptr, A.indices, A.data, coarsening="pmis", min_size=1000, symmetrize=False)
    L = csr(h.fine_ichol(), (120, 120))
    b = np.random.default_rng(3).random(120)
    x = np.zeros(120)
    lib.h_ichol_solve(h.h, 0, dp(b), dp(x))
    y = spla.spsolve_triangular(L.tocsr(), b, lower=True)
    xr = spla.spsolve_triangular(L.T.tocsr(), y, lower=False)
    assert np.allclose(x, xr, rtol=1e-12)
    # SpMV / residual / transpose product against scipy
    xv = np.random.default_rng(4).random(120)
    y1 = np.zeros(120)
    lib.abyvg(ip(A.indptr), ip(A.indices), dp(A.data), dp(xv), 120, dp(y1))
    assert np.allclose(y1, A @ xv, rtol=1e-14)
    y2 = np.zeros(120)
    lib.abyvam(dp(b), ip(A.indptr), ip(A.indices), dp(A.data), dp(xv), 120, dp(y2))
    assert np.allclose(y2, b - A @ xv, rtol=1e-13, atol=1e-13)
    B = sp.random(120, 40, density=0.1, random_state=1, format="csr")
    y3 = np.zeros(40)
    lib.vbya(ip(B.indptr), ip(B.indices), dp(B.data), dp(xv), 120, 40, dp(y3))
    assert np.allclose(y3, B.T @ xv, rtol=1e-13)


def test_pcg_converges_to_direct_solution(oracle):
    net = make_network(16, 9)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="rs")
    x, it, _ = h.solve(net.b, 1e-10, 1000)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    xd = spla.spsolve(A.tocsc(), net.b)
    assert 0 < it < 50
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


def test_preconditioner_is_symmetric(oracle):
    # B = amg_ichol_precond is a symmetric operator: u'B v == v'B u
    net = make_network(12, 6, seed=4)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=300)
    rng = np.random.default_rng(7)
    u, v = rng.random(net.n), rng.random(net.n)
    a, b = u @ h.precond(v), v @ h.precond(u)
    assert abs(a - b) <= 1e-6 * abs(a)


def _sweep_closed_form(A, b, x0, order, nu, direction):
    # nu sequential sweeps in `order` (forward) or its reverse (backward):
    # with B = A[pi, pi] = D + L + U in visiting order, a forward sweep
    # solves (D + L) x'[pi] = b[pi] - U x[pi], a backward one
    # (D + U) x'[pi] = b[pi] - L x[pi]
    Ap = A.toarray()[np.ix_(order, order)]
    x = x0[order].copy()
    for _ in range(nu):
        if direction == 0:
            x = np.linalg.solve(np.tril(Ap), b[order] - np.triu(Ap, 1) @ x)
        else:
            x = np.linalg.solve(np.triu(Ap), b[order] - np.tril(Ap, -1) @ x)
    out = np.empty_like(x)
    out[order] = x
    return out


@pytest.mark.parametrize("nu,direction", [(1, 0), (1, 1), (2, 0), (3, 1)])
def test_mex_smoother_restatement_closed_form(oracle, nu, direction):
    """oracle_cf_gauss_seidel (Matlab_amg/cf_gauss_seidel_mex.c:27-105:
    1-based cf_order in doubles, direction 0 forward / 1 backward, nu
    sweeps) against the dense closed form of the sequential sweep; a
    dropped term, a wrong diagonal or a reversed direction fails it."""
    import ctypes as C
    rng = np.random.default_rng(20 + nu + direction)
    n = 70
    M = sp.random(n, n, density=0.07, random_state=nu + 10 * direction)
    A = (M + M.T).tocsr()
    A.setdiag(np.abs(A).sum(axis=1).A.ravel() + 0.5)
    A.sort_indices()
    ia, ja, va = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    order = rng.permutation(n)
    b, x0 = rng.random(n), rng.random(n)
    f = oracle.lib.oracle_cf_gauss_seidel
    f.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double)]
    x = np.zeros(n)
    o1 = order.astype(np.float64) + 1.0
    f(n, ip(ia), ip(ja), dp(va), dp(b), dp(x0), dp(o1), nu, direction, dp(x))
    ref = _sweep_closed_form(A, b, x0, order, nu, direction)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-13
