"""Host-side schedules the device kernels rely on: row blocks for the
adaptive CSR kernel, multicolour orders and triangular level sets.  Each
is checked against the property the kernel needs, not against a second
copy of the construction."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import _lib as L
from blood_network_amd import make_network

lib = L.lib
NNZ, ROWS = 2048, 256


def row_blocks(ia, max_nnz=NNZ, max_rows=ROWS):
    n = len(ia) - 1
    out = np.zeros(n + 1, np.int32)
    nb = lib.bn_row_blocks(L.ip(ia), n, max_nnz, max_rows, L.ip(out))
    return out[: nb + 1]


def check_blocks(ia, blk, max_nnz=NNZ, max_rows=ROWS):
    n = len(ia) - 1
    lens = np.diff(ia)
    assert blk[0] == 0 and blk[-1] == n
    assert np.all(np.diff(blk) >= 1)
    for a, b in zip(blk[:-1], blk[1:]):
        tot = int(ia[b] - ia[a])
        if tot > max_nnz:
            # vector mode: exactly one over-long row
            assert b == a + 1 and lens[a] > max_nnz
            continue
        assert b - a <= max_rows
        # greedy and maximal: the next row would break a limit (or is long)
        if b < n:
            assert b - a == max_rows or tot + lens[b] > max_nnz


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_row_blocks_cover_and_respect_limits(seed):
    rng = np.random.default_rng(seed)
    n = 5000
    lens = rng.integers(0, 12, n)
    lens[rng.choice(n, 30, replace=False)] = rng.integers(300, 6000, 30)  # hubs, some > 2048
    lens[rng.choice(n, 200, replace=False)] = 0                           # empty rows
    ia = np.zeros(n + 1, np.int32)
    ia[1:] = np.cumsum(lens)
    blk = row_blocks(ia)
    check_blocks(ia, blk)
    assert np.sum(np.diff(blk) == 1) >= np.sum(lens > NNZ)


def test_row_blocks_edge_cases():
    # all rows empty: blocks of max_rows rows
    ia = np.zeros(1001, np.int32)
    blk = row_blocks(ia)
    check_blocks(ia, blk)
    assert len(blk) - 1 == 4
    # one row exactly at the limit, one just over it
    ia = np.array([0, NNZ, 2 * NNZ + 1], np.int32)
    blk = row_blocks(ia)
    assert list(blk) == [0, 1, 2]
    # single row
    ia = np.array([0, 3], np.int32)
    assert list(row_blocks(ia)) == [0, 1]


def test_row_blocks_on_a_network_level(net16):
    blk = row_blocks(net16.ia)
    check_blocks(net16.ia, blk)


def test_colour_order_is_a_proper_colouring():
    net = make_network(12, 6, seed=5)
    n = net.n
    m = L.CSR(n, n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(net.val))
    mis = (np.arange(n) % 4 == 0).astype(np.int32)
    perm = np.zeros(n, np.int32)
    cp = C.POINTER(C.c_int)()
    ncol = lib.bn_colour_order(C.byref(m), L.ip(mis), 1, L.ip(perm), C.byref(cp))
    col_ptr = np.ctypeslib.as_array(cp, (ncol + 1,)).copy()
    assert sorted(perm) == list(range(n))
    colour = np.empty(n, np.int64)
    for c in range(ncol):
        colour[perm[col_ptr[c]:col_ptr[c + 1]]] = c
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(n, n)).tocoo()
    off = A.row != A.col
    assert not np.any(colour[A.row[off]] == colour[A.col[off]])
    # C points first, each colour ascending in the old numbering
    nC = int(mis.sum())
    assert np.all(mis[perm[:nC]] == 1)
    for c in range(ncol):
        seg = perm[col_ptr[c]:col_ptr[c + 1]]
        assert np.all(np.diff(seg) > 0)


def colour_of(ia, ja, val, mis, split):
    n = len(ia) - 1
    m = L.CSR(n, n, len(ja), L.ip(ia), L.ip(ja), L.dp(val))
    perm = np.zeros(n, np.int32)
    cp = C.POINTER(C.c_int)()
    ncol = lib.bn_colour_order(C.byref(m), L.ip(mis), split, L.ip(perm), C.byref(cp))
    col_ptr = np.ctypeslib.as_array(cp, (ncol + 1,)).copy()
    colour = np.empty(n, np.int64)
    for c in range(ncol):
        colour[perm[col_ptr[c]:col_ptr[c + 1]]] = c
    A = sp.csr_matrix((val, ja, ia), shape=(n, n)).tocoo()
    off = A.row != A.col
    assert not np.any(colour[A.row[off]] == colour[A.col[off]])   # a proper colouring
    return ncol, colour


def grid(k, diag=False):
    n = k * k
    rows, cols = [], []
    for i in range(k):
        for j in range(k):
            u = i * k + j
            for di, dj in ((0, 1), (1, 0)) + (((1, 1),) if diag else ()):
                if i + di < k and j + dj < k:
                    v = (i + di) * k + j + dj
                    rows += [u, v]
                    cols += [v, u]
    A = sp.csr_matrix((-np.ones(len(rows)), (rows, cols)), shape=(n, n)) + 4.0 * sp.identity(n)
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def test_bipartite_couplings_take_two_colours():
    # a grid has no odd cycle: two colours by parity (the greedy pass in row
    # order also finds two here; the parity colouring is the checkerboard)
    ia, ja, val = grid(9)
    n = len(ia) - 1
    ncol, colour = colour_of(ia, ja, val, np.zeros(n, np.int32), 0)
    i, j = np.divmod(np.arange(n), 9)
    assert ncol == 2 and np.array_equal(colour, (i + j) % 2)
    # the C set of the checkerboard couples to nothing of its own: one
    # colour per class
    mis = ((i + j) % 2 == 0).astype(np.int32)
    ncol, colour = colour_of(ia, ja, val, mis, 1)
    assert ncol == 2 and np.array_equal(colour, 1 - mis)


def test_crown_graph_takes_two_colours_where_greedy_takes_three():
    # a_i couples to b_j for i != j, rows ordered a1 b1 a2 b2 a3 b3: the
    # greedy pass in row order needs three colours, the graph is bipartite
    # and the parity colouring takes two (the a's and the b's)
    k = 3
    edges = [(2 * i, 2 * j + 1) for i in range(k) for j in range(k) if i != j]
    n = 2 * k
    rows = [a for a, b in edges] + [b for a, b in edges]
    cols = [b for a, b in edges] + [a for a, b in edges]
    A = (sp.csr_matrix((-np.ones(len(rows)), (rows, cols)), shape=(n, n)) + 4.0 * sp.identity(n)).tocsr()
    A.sort_indices()
    ia, ja, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    ncol, colour = colour_of(ia, ja, val, np.zeros(n, np.int32), 0)
    assert ncol == 2 and np.array_equal(colour, np.arange(n) % 2)


def test_odd_cycle_falls_back_to_the_greedy_colouring():
    # diagonals make triangles: no two-colouring exists, the greedy pass
    # colours (properly, checked in colour_of) with three or more
    ia, ja, val = grid(7, diag=True)
    n = len(ia) - 1
    ncol, _ = colour_of(ia, ja, val, np.zeros(n, np.int32), 0)
    assert ncol >= 3


def test_level_sets_respect_dependencies():
    net = make_network(10, 5, seed=6)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    Lo = sp.tril(A).tocsr()
    Lo.sort_indices()
    n = net.n
    rows = np.zeros(n, np.int32)
    pp = C.POINTER(C.c_int)()
    ia, ja = Lo.indptr.astype(np.int32), Lo.indices.astype(np.int32)
    nl = lib.bn_level_sets(L.ip(ia), L.ip(ja), n, 1, L.ip(rows), C.byref(pp))
    ptr = np.ctypeslib.as_array(pp, (nl + 1,)).copy()
    level = np.empty(n, np.int64)
    for k in range(nl):
        level[rows[ptr[k]:ptr[k + 1]]] = k
    assert sorted(rows) == list(range(n))
    coo = Lo.tocoo()
    dep = coo.col < coo.row
    # every dependency sits in an earlier level, and levels are tight
    assert np.all(level[coo.col[dep]] < level[coo.row[dep]])
    need = np.zeros(n, np.int64)
    np.maximum.at(need, coo.row[dep], level[coo.col[dep]] + 1)
    assert np.array_equal(level, need)


def chunk_order(ia, ja, val, mis, nchunks, split=1):
    n = len(ia) - 1
    m = L.CSR(n, n, len(ja), L.ip(ia), L.ip(ja), L.dp(val))
    perm = np.zeros(n, np.int32)
    cp, cc, cb = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    nk = lib.bn_chunk_order(C.byref(m), L.ip(mis) if mis is not None else None, split, nchunks, L.ip(perm),
                            C.byref(cp), C.byref(cc), C.byref(cb))
    chunk = np.ctypeslib.as_array(cp, (nk + 1,)).copy()
    ccol = np.ctypeslib.as_array(cc, (nk + 1,)).copy()
    cbnd = np.ctypeslib.as_array(cb, (int(ccol[-1]),)).copy()
    return nk, perm, chunk, ccol, cbnd


@pytest.mark.parametrize("nchunks", [1, 7, 64, 10**6])
def test_chunk_order_colours_each_chunk_properly(nchunks):
    net = make_network(12, 6, seed=7)
    n = net.n
    mis = (np.arange(n) % 3 == 0).astype(np.int32)
    nk, perm, chunk, ccol, cbnd = chunk_order(net.ia, net.ja, net.val, mis, nchunks)
    S = -(-n // min(nchunks, n))
    assert nk == -(-n // S) and chunk[0] == 0 and chunk[-1] == n
    assert np.array_equal(chunk[:-1], np.arange(nk) * S)
    assert sorted(perm) == list(range(n))
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(n, n)).tocoo()
    colour = np.full(n, -1)
    chunk_of = np.empty(n, np.int64)
    for k in range(nk):
        lo, hi = chunk[k], chunk[k + 1]
        seg = perm[lo:hi]
        # the permutation never leaves the chunk
        assert seg.min() >= lo and seg.max() < hi
        b = cbnd[ccol[k]:ccol[k + 1]]
        assert b[0] == lo and b[-1] == hi and np.all(np.diff(b) > 0)
        for c in range(len(b) - 1):
            rows = perm[b[c]:b[c + 1]]
            assert np.all(np.diff(rows) > 0)          # ascending within a colour
            colour[rows] = c
            # a colour is all-C or all-F, C colours first
            assert len(set(mis[rows])) == 1
        cls = mis[perm[lo:hi]]
        assert np.all(np.diff(-cls) >= 0)             # C rows (1) before F rows (0)
        chunk_of[lo:hi] = k
    # no coupling inside a chunk joins two rows of one colour
    off = A.row != A.col
    same_chunk = chunk_of[A.row] == chunk_of[A.col]
    sel = off & same_chunk
    assert not np.any(colour[A.row[sel]] == colour[A.col[sel]])



def sweep_levels(A, order, backward):
    """bn_sweep_levels on a scipy CSR matrix: level of each row."""
    n = A.shape[0]
    ia, ja = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    va = A.data.astype(np.float64)
    m = L.CSR(n, n, len(ja), L.ip(ia), L.ip(ja), L.dp(va))
    rows = np.zeros(n, np.int32)
    pp = C.POINTER(C.c_int)()
    o = None if order is None else np.ascontiguousarray(order, np.int32)
    nl = lib.bn_sweep_levels(C.byref(m), L.ip(o) if o is not None else None, int(backward), L.ip(rows),
                             C.byref(pp))
    if nl < 0:
        return None
    ptr = np.ctypeslib.as_array(pp, (nl + 1,)).copy()
    level = np.empty(n, np.int64)
    for k in range(nl):
        level[rows[ptr[k]:ptr[k + 1]]] = k
        assert np.all(np.diff(rows[ptr[k]:ptr[k + 1]]) > 0)
    assert sorted(rows) == list(range(n))
    return level


def sequential_sweep(A, b, x, order):
    """The reference's sweep (C_amg1/src/amg_cycle.c:21-80) row by row."""
    x = x.copy()
    for i in order:
        s, d = b[i], 1.0
        for k in range(A.indptr[i], A.indptr[i + 1]):
            j = A.indices[k]
            if j == i:
                d = A.data[k]
            else:
                s -= A.data[k] * x[j]
        x[i] = s / d
    return x


def levelled_sweep(A, b, x, level):
    """The same row updates, all rows of a level reading the iterate as it
    stood before the level (what one device launch per level does)."""
    x = x.copy()
    for lv in range(level.max() + 1):
        rows = np.flatnonzero(level == lv)
        new = {}
        for i in rows:
            s, d = b[i], 1.0
            for k in range(A.indptr[i], A.indptr[i + 1]):
                j = A.indices[k]
                if j == i:
                    d = A.data[k]
                else:
                    s -= A.data[k] * x[j]
            new[i] = s / d
        for i, v in new.items():
            x[i] = v
    return x


@pytest.mark.parametrize("kind", ["natural", "cf", "random"])
@pytest.mark.parametrize("backward", [0, 1])
def test_sweep_levels_reproduce_the_sequential_sweep(kind, backward):
    net = make_network(8, 5, seed=9)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    n = net.n
    rng = np.random.default_rng(3)
    if kind == "natural":
        order = None
        seq = np.arange(n)
    elif kind == "cf":
        c = rng.random(n) < 0.3                      # C points first, then F, ascending
        seq = np.concatenate([np.flatnonzero(c), np.flatnonzero(~c)])
        order = seq
    else:
        seq = rng.permutation(n)
        order = seq
    level = sweep_levels(A, order, backward)
    visit = seq[::-1] if backward else seq
    pos = np.empty(n, np.int64)
    pos[visit] = np.arange(n)
    # a coupled pair sits in levels ordered as the sweep visits them, and
    # levels are tight: one more than the latest earlier-visited neighbour
    coo = (A + A.T).tocoo()
    off = coo.row != coo.col
    r, c = coo.row[off], coo.col[off]
    before = pos[c] < pos[r]
    assert np.all(level[c[before]] < level[r[before]])
    need = np.zeros(n, np.int64)
    np.maximum.at(need, r[before], level[c[before]] + 1)
    assert np.array_equal(level, need)
    # and the level-by-level update gives the sequential sweep bit for bit
    b = rng.random(n)
    x0 = rng.random(n)
    assert np.array_equal(levelled_sweep(A, b, x0, level), sequential_sweep(A, b, x0, visit))


def test_sweep_levels_couple_through_either_triangle():
    # a pattern stored in one triangle only: row 2 reads x_0, row 0 does not
    # store a_02; visiting 2 before 0 must still separate them
    A = sp.csr_matrix(np.array([[4.0, 0, 0], [0, 4.0, 0], [1.0, 0, 4.0]]))
    level = sweep_levels(A, np.array([2, 0, 1]), 0)
    assert level[2] < level[0]
    level = sweep_levels(A, None, 0)
    assert level[0] < level[2]


def test_sweep_levels_reject_a_non_permutation():
    A = sp.identity(4, format="csr")
    assert sweep_levels(A, np.array([0, 1, 1, 3]), 0) is None
    assert sweep_levels(A, np.array([0, 1, 2, 4]), 0) is None
