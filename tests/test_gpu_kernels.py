"""Kernel-level parity: each HIP kernel, called through the C ABI on device
buffers, against the oracle (or against the closed form the oracle's
routine computes).  FP64 bar: 1e-12 relative (summation order differs from
the sequential CPU loops); FP32 bar: 1e-5 relative."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import AMGSolver, make_network
from blood_network_amd import _lib as L
from oracle_lib import dp, ip

pytestmark = pytest.mark.gpu
lib = L.lib


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def dev(a):
    return L.DeviceArray(np.ascontiguousarray(a))


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def coarse_level_matrix():
    """A Galerkin level with long rows (tree hubs) plus its P."""
    net = make_network(24, 12, seed=2)
    s = AMGSolver(net.ia, net.ja, net.val, coarsening="pmis", min_size=200)
    return s, net


@pytest.mark.parametrize("vt", ["d", "s"])
def test_csr_ops_against_oracle(oracle, vt):
    s, net = coarse_level_matrix()
    rng = np.random.default_rng(0)
    dt = np.float64 if vt == "d" else np.float32
    tol = 1e-12 if vt == "d" else 1e-5
    for lev in range(s.num_levels()):
        ia, ja, val = s.level(lev)["A"]
        n = len(ia) - 1
        lens = np.diff(ia)
        for lanes in (1, 4, 16, 64, 1 | L.LANES_DEEP8, 1 | L.LANES_DEEP16, 2 | L.LANES_DEEP16, 4 | L.LANES_DEEP8,
                      8 | L.LANES_DEEP16, 16 | L.LANES_DEEP8, 16 | L.LANES_DEEP16, 64 | L.LANES_DEEP16):
            long_len = 16 * (lanes & 0xFF)
            long_rows = np.where(lens > long_len)[0].astype(np.int32)
            case(f"csr level {lev} n={n} lanes={lanes & 0xFF} rounds={lanes >> 8} long={len(long_rows)} vt={vt}")
            x = rng.random(n)
            b = rng.random(n)
            y0 = rng.random(n)
            ref_set = np.zeros(n)
            oracle.lib.abyvg(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_set))
            ref_res = np.zeros(n)
            oracle.lib.abyvam(dp(b), ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_res))
            ref_add = y0.copy()
            oracle.lib.abyvcs(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_add))
            dia, dja, dv = dev(ia), dev(ja), dev(val.astype(dt))
            dx, db = dev(x.astype(dt)), dev(b.astype(dt))
            dlr = dev(long_rows if len(long_rows) else np.zeros(1, np.int32))
            for op, ref, init in ((0, ref_set, None), (1, ref_res, None), (2, ref_add, y0)):
                dy = dev((init if init is not None else np.zeros(n)).astype(dt))
                lib.bn_csr_apply_binned(vt.encode(), op, n, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy.ptr, db.ptr,
                                        lanes, long_len, len(long_rows), dlr.ptr, None)
                got = dy.numpy().astype(np.float64)
                assert relerr(got, ref) < tol, (lev, lanes, op)


def hub_matrix(seed=0):
    """Square CSR with empty rows, short rows and hub rows longer than one
    row block (vector mode), columns sorted and unique per row."""
    rng = np.random.default_rng(seed)
    n = 6000
    lens = rng.integers(1, 14, n)
    lens[rng.choice(n, 300, replace=False)] = 0
    lens[rng.choice(n, 12, replace=False)] = rng.integers(2049, 5900, 12)
    lens[rng.choice(n, 40, replace=False)] = rng.integers(200, 2048, 40)
    ia = np.zeros(n + 1, np.int32)
    ia[1:] = np.cumsum(lens)
    ja = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    val = rng.random(len(ja)) + 0.5
    return ia, ja, val


@pytest.mark.parametrize("vt", ["d", "s"])
def test_row_block_kernel_against_oracle(oracle, vt):
    s, net = coarse_level_matrix()
    mats = [("A%d" % lev, s.level(lev)["A"]) for lev in range(s.num_levels())]
    mats.append(("hubs", hub_matrix()))
    rng = np.random.default_rng(8)
    dt = np.float64 if vt == "d" else np.float32
    tol = 1e-12 if vt == "d" else 1e-5
    for name, (ia, ja, val) in mats:
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        n = len(ia) - 1
        blk = np.zeros(n + 1, np.int32)
        nb = lib.bn_row_blocks(ip(ia), n, 2048, 256, ip(blk))
        case(f"row-block csr {name} n={n} nnz={len(ja)} blocks={nb} vt={vt}")
        x, b, y0 = rng.random(n), rng.random(n), rng.random(n)
        ref_set, ref_res, ref_add = np.zeros(n), np.zeros(n), y0.copy()
        oracle.lib.abyvg(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_set))
        oracle.lib.abyvam(dp(b), ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_res))
        oracle.lib.abyvcs(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_add))
        dia, dja, dv, dblk = dev(ia), dev(ja), dev(val.astype(dt)), dev(blk[: nb + 1])
        dblkq = dev(ia[blk[: nb + 1]].astype(np.int32))      # entry offsets of the block boundaries
        dx, db = dev(x.astype(dt)), dev(b.astype(dt))
        bits = np.uint64 if vt == "d" else np.uint32
        for op, ref, init in ((0, ref_set, None), (1, ref_res, None), (2, ref_add, y0)):
            outs = []
            for q in (None, dblkq.ptr):
                dy = dev((init if init is not None else np.zeros(n)).astype(dt))
                lib.bn_csr_apply_blocks(vt.encode(), op, nb, dblk.ptr, q, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy.ptr,
                                        db.ptr, None)
                outs.append(dy.numpy())
                got = outs[-1].astype(np.float64)
                assert relerr(got, ref) < tol, (name, op)
            # with or without the stored entry offsets: the same sums
            assert np.array_equal(outs[0].view(bits), outs[1].view(bits)), (name, op)
        # the scaled residual writes the plain residual bit for bit and its
        # product with dsc (hub rows take the one-long-row path)
        dsc = (rng.random(n) + 0.5).astype(dt)
        dy1, dy2, dys, dd = dev(np.zeros(n, dt)), dev(np.zeros(n, dt)), dev(np.zeros(n, dt)), dev(dsc)
        lib.bn_csr_apply_blocks(vt.encode(), 1, nb, dblk.ptr, None, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy1.ptr, db.ptr,
                                None)
        lib.bn_csr_resid_blocks_scaled(vt.encode(), nb, dblk.ptr, dblkq.ptr, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy2.ptr,
                                       db.ptr, dd.ptr, dys.ptr, None)
        g1, g2 = dy1.numpy(), dy2.numpy()
        assert np.array_equal(g1.view(bits), g2.view(bits)), name
        assert np.array_equal(dys.numpy().view(bits), (g2 * dsc).astype(dt).view(bits)), name


def test_row_block_product_with_fused_dot():
    """q = A p with p.q fused: q bit for bit the plain row-block product's,
    one partial per row block (the block's w_r y_r, hub rows by the
    one-long-row path), and the finish over those slots giving p'Ap and
    alpha = r'z / p'Ap."""
    s, net = coarse_level_matrix()
    mats = [("A%d" % lev, s.level(lev)["A"]) for lev in range(s.num_levels())]
    mats.append(("hubs", hub_matrix()))
    rng = np.random.default_rng(12)
    for name, (ia, ja, val) in mats:
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        n = len(ia) - 1
        blk = np.zeros(n + 1, np.int32)
        nb = lib.bn_row_blocks(ip(ia), n, 2048, 256, ip(blk))
        case(f"row-block csr with fused dot {name} n={n} nnz={len(ja)} blocks={nb}")
        x, w = rng.random(n) - 0.5, rng.random(n) - 0.5
        dia, dja, dv, dblk = dev(ia), dev(ja), dev(np.asarray(val, np.float64)), dev(blk[: nb + 1])
        dblkq = dev(ia[blk[: nb + 1]].astype(np.int32))
        dx, dw = dev(x), dev(w)
        dy0, dy1 = dev(np.zeros(n)), dev(np.zeros(n))
        dpart = dev(np.full(nb, np.nan))
        lib.bn_csr_apply_blocks(b"d", 0, nb, dblk.ptr, dblkq.ptr, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy0.ptr, None,
                                None)
        lib.bn_csr_apply_blocks_dot(nb, dblk.ptr, dblkq.ptr, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy1.ptr, dw.ptr,
                                    dpart.ptr, None)
        y0, y1, part = dy0.numpy(), dy1.numpy(), dpart.numpy()
        assert np.array_equal(y0.view(np.uint64), y1.view(np.uint64)), name
        terms = w * y1
        want = np.array([terms[blk[k]:blk[k + 1]].sum() for k in range(nb)])
        scale = np.array([np.abs(terms[blk[k]:blk[k + 1]]).sum() for k in range(nb)]) + 1e-300
        assert np.all(np.abs(part - want) <= 1e-13 * scale), name
        sc = np.zeros(16)
        sc[0] = 3.25                                     # r'z
        dsc = dev(sc)
        lib.bn_scalar_finish_pap(dpart.ptr, nb, dsc.ptr, None)
        got = dsc.numpy()
        pap = terms.sum()
        assert abs(got[2] - pap) <= 1e-13 * np.abs(terms).sum(), name
        assert abs(got[3] - 3.25 / got[2]) <= 1e-15 * abs(got[3]), name


# the sentinel of values not written yet (a signalling-NaN bit pattern)
SENTINEL64 = np.uint64(0x7ff4dead0000beef)


def seg_bin(ia, long_len, seg=1024):
    """Long-row bin with segments: rows longer than long_len, longest first
    (ties by index), cut into `seg`-entry segments."""
    lens = np.diff(ia)
    rows = np.where(lens > long_len)[0]
    rows = rows[np.lexsort((rows, -lens[rows]))].astype(np.int32)
    nseg_row = -(-lens[rows] // seg)
    seg_ptr = np.zeros(len(rows) + 1, np.int32)
    seg_ptr[1:] = np.cumsum(nseg_row)
    seg_row = np.repeat(np.arange(len(rows), dtype=np.int32), nseg_row)
    nmulti = int(np.sum(lens[rows] > seg))
    return rows, seg_ptr, seg_row, nmulti


@pytest.mark.parametrize("vt", ["d", "s"])
def test_segmented_long_rows_against_oracle(oracle, vt):
    s, net = coarse_level_matrix()
    mats = [("A%d" % lev, s.level(lev)["A"]) for lev in range(s.num_levels())]
    mats.append(("hubs", hub_matrix(1)))
    rng = np.random.default_rng(9)
    dt = np.float64 if vt == "d" else np.float32
    tol = 1e-12 if vt == "d" else 1e-5
    for name, (ia, ja, val) in mats:
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        n = len(ia) - 1
        for lanes in (1, 4, 16, 16 | L.LANES_DEEP16):
            rows, seg_ptr, seg_row, nmulti = seg_bin(ia, 16 * (lanes & 0xFF))
            case(f"segmented csr {name} n={n} lanes={lanes & 0xFF} rounds={lanes >> 8} long={len(rows)} "
                 f"segments={seg_ptr[-1]} multi={nmulti}")
            x, b, y0 = rng.random(n), rng.random(n), rng.random(n)
            ref_set, ref_res, ref_add = np.zeros(n), np.zeros(n), y0.copy()
            oracle.lib.abyvg(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_set))
            oracle.lib.abyvam(dp(b), ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_res))
            oracle.lib.abyvcs(ip(ia), ip(ja), dp(val), dp(x), n, dp(ref_add))
            dia, dja, dv = dev(ia), dev(ja), dev(val.astype(dt))
            dx, db = dev(x.astype(dt)), dev(b.astype(dt))
            one = np.zeros(1, np.int32)
            drows = dev(rows if len(rows) else one)
            dsp, dsr = dev(seg_ptr), dev(seg_row if len(seg_row) else one)
            # the multi-segment rows' partials pass through part (the
            # sentinel between calls) and are counted in cnt (zero between
            # calls): every call must leave both as it found them
            nslot = int(seg_ptr[nmulti])
            dpart = dev(np.zeros(max(1, nslot)))
            dcnt = dev(np.ones(max(1, nmulti), np.int32))
            lib.bn_seg_reset(nslot, dpart.ptr, nmulti, dcnt.ptr, None)
            for op, ref, init in ((0, ref_set, None), (1, ref_res, None), (2, ref_add, y0)):
                dy = dev((init if init is not None else np.zeros(n)).astype(dt))
                lib.bn_csr_apply_seg(vt.encode(), op, n, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy.ptr, db.ptr, lanes,
                                     16 * (lanes & 0xFF), len(rows), drows.ptr, dsp.ptr, dsr.ptr, int(seg_ptr[-1]),
                                     dpart.ptr, dcnt.ptr, None)
                got = dy.numpy().astype(np.float64)
                assert relerr(got, ref) < tol, (name, lanes, op)
                assert np.all(dpart.numpy()[:nslot].view(np.uint64) == SENTINEL64), (name, lanes, op)
                assert np.all(dcnt.numpy()[:nmulti] == 0), (name, lanes, op)


def test_restriction_gather_equals_scatter(oracle):
    s, net = coarse_level_matrix()
    rng = np.random.default_rng(1)
    for lev in range(s.num_levels() - 1):
        lv = s.level(lev)
        pia, pja, pv = lv["P"]
        n, nc = len(pia) - 1, int(lv["mis"].sum())
        case(f"restriction level {lev} {n}->{nc}")
        r = rng.random(n)
        ref = np.zeros(nc)
        oracle.lib.vbya(ip(pia), ip(pja), dp(pv), dp(r), n, nc, dp(ref))
        R = sp.csr_matrix((pv, pja, pia), shape=(n, nc)).T.tocsr()
        R.sort_indices()
        dia, dja, dv = dev(R.indptr.astype(np.int32)), dev(R.indices.astype(np.int32)), dev(R.data)
        dx, dy = dev(r), dev(np.zeros(nc))
        lib.bn_csr_apply(b"d", 0, nc, dia.ptr, dja.ptr, dv.ptr, dx.ptr, dy.ptr, None, 8, None, None, None)
        assert relerr(dy.numpy(), ref) < 1e-12


@pytest.mark.parametrize("vt", ["d", "s"])
def test_multicolour_sweep_is_exact_gauss_seidel(vt):
    # A forward sweep in order pi solves (D + L_pi) x_new = b - U_pi x_old
    # exactly (C_amg1/src/amg_cycle.c:52-65 in that order); the colour
    # kernels must produce that x_new.
    net = make_network(14, 8, seed=3)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    n = net.n
    m = L.CSR(n, n, net.nnz, ip(net.ia), ip(net.ja), dp(net.val))
    perm = np.zeros(n, np.int32)
    cptr = C.POINTER(C.c_int)()
    mis = (np.arange(n) % 3 == 0).astype(np.int32)  # any C/F split works
    ncol = lib.bn_colour_order(C.byref(m), ip(mis), 1, ip(perm), C.byref(cptr))
    col_ptr = np.ctypeslib.as_array(cptr, (ncol + 1,)).copy()
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    Ap = A[perm][:, perm].tocsr()
    Ap.sort_indices()
    case(f"multicolour GS n={n} colours={ncol} vt={vt}")
    rng = np.random.default_rng(4)
    b = rng.random(n) * 1e-9
    x_old = rng.random(n) * 1e3
    D = Ap.diagonal()
    dt = np.float64 if vt == "d" else np.float32
    dia, dja, dv = dev(Ap.indptr.astype(np.int32)), dev(Ap.indices.astype(np.int32)), dev(Ap.data.astype(dt))
    ddinv, db, dx = dev((1.0 / D).astype(dt)), dev(b.astype(dt)), dev(x_old.astype(dt))
    for c in range(ncol):
        lib.bn_gs_color(vt.encode(), int(col_ptr[c]), int(col_ptr[c + 1]), dia.ptr, dja.ptr, dv.ptr, ddinv.ptr,
                        db.ptr, dx.ptr, 2, None, None, None)
    x_new = dx.numpy().astype(np.float64)
    Lp = sp.tril(Ap, k=0).tocsr()
    Up = sp.triu(Ap, k=1).tocsr()
    lhs = Lp @ x_new
    rhs = b - Up @ x_old
    assert relerr(lhs, rhs) < (1e-12 if vt == "d" else 2e-5)


@pytest.mark.parametrize("vt", ["d", "s"])
@pytest.mark.parametrize("lanes", [1, 4])
def test_mapped_multicolour_sweep_is_exact_gauss_seidel(vt, lanes):
    # The solver's layout: vectors in natural order, the level's rows
    # copied in colour order (columns natural), row r updating entry
    # map[r].  A forward sweep in colour order pi solves
    # (D + L_pi) x_new = b - U_pi x_old, L_pi/U_pi = couplings to rows
    # earlier/later in pi.
    net = make_network(14, 8, seed=3)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    n = net.n
    m = L.CSR(n, n, net.nnz, ip(net.ia), ip(net.ja), dp(net.val))
    perm = np.zeros(n, np.int32)
    cptr = C.POINTER(C.c_int)()
    mis = (np.arange(n) % 3 == 0).astype(np.int32)
    ncol = lib.bn_colour_order(C.byref(m), ip(mis), 1, ip(perm), C.byref(cptr))
    col_ptr = np.ctypeslib.as_array(cptr, (ncol + 1,)).copy()
    Ac = A[perm].tocsr()                      # rows in colour order, columns natural
    Ac.sort_indices()
    pos = np.empty(n, np.int64)
    pos[perm] = np.arange(n)
    case(f"mapped multicolour GS n={n} colours={ncol} vt={vt} lanes={lanes}")
    rng = np.random.default_rng(4)
    b = rng.random(n) * 1e-9
    x_old = rng.random(n) * 1e3
    zbase = rng.random(n)
    D = A.diagonal()
    dt = np.float64 if vt == "d" else np.float32
    dia, dja, dv = dev(Ac.indptr.astype(np.int32)), dev(Ac.indices.astype(np.int32)), dev(Ac.data.astype(dt))
    dmap, ddinv = dev(perm), dev((1.0 / D[perm]).astype(dt))
    db, dx = dev(b.astype(dt)), dev(x_old.astype(dt))
    dzb, dzs = dev(zbase.astype(dt)), dev(np.zeros(n, dt))
    for c in range(ncol):
        lib.bn_gs_color_map(vt.encode(), int(col_ptr[c]), int(col_ptr[c + 1]), dmap.ptr, dia.ptr, dja.ptr, dv.ptr,
                            ddinv.ptr, db.ptr, dx.ptr, lanes, dzb.ptr, dzs.ptr, None)
    x_new = dx.numpy().astype(np.float64)
    coo = A.tocoo()
    lower = pos[coo.col] < pos[coo.row]
    upper = pos[coo.col] > pos[coo.row]
    Lp = sp.csr_matrix((coo.data[lower], (coo.row[lower], coo.col[lower])), shape=(n, n))
    Up = sp.csr_matrix((coo.data[upper], (coo.row[upper], coo.col[upper])), shape=(n, n))
    lhs = D * x_new + Lp @ x_new
    rhs = b - Up @ x_old
    assert relerr(lhs, rhs) < (1e-12 if vt == "d" else 2e-5)
    # every colour wrote zsum = zbase + x for its rows
    assert relerr(dzs.numpy().astype(np.float64), zbase + x_new) < (1e-15 if vt == "d" else 1e-6)
    # zero start: the first colour is a pure diagonal scaling, and the
    # same pass zeroes every later colour's rows
    dx0 = dev(np.full(n, np.nan, dt))
    lo, hi = int(col_ptr[0]), int(col_ptr[1])
    lib.bn_gs_color_zero_map(vt.encode(), lo, hi, n, dmap.ptr, ddinv.ptr, db.ptr, dx0.ptr, None)
    got = dx0.numpy().astype(np.float64)
    assert relerr(got[perm[lo:hi]], (b / D)[perm[lo:hi]]) < (1e-15 if vt == "d" else 1e-6)
    assert np.all(got[perm[hi:]] == 0.0)
    # the same step in level-row order, from a one-byte colour index and
    # 1/a_ii in level-row order: bit for bit the colour-order form's
    # values, +0.0 on every other row, for the first and the last colour
    colour = np.empty(n, np.uint8)
    for c in range(ncol):
        colour[perm[col_ptr[c]:col_ptr[c + 1]]] = c
    dcol, ddn = dev(colour), dev((1.0 / D).astype(dt))
    bits = np.uint64 if vt == "d" else np.uint32
    for c in (0, ncol - 1):
        lo, hi = int(col_ptr[c]), int(col_ptr[c + 1])
        dm, dcl = dev(np.full(n, np.nan, dt)), dev(np.full(n, np.nan, dt))
        lib.bn_gs_color_zero_map(vt.encode(), lo, hi, n, dmap.ptr, ddinv.ptr, db.ptr, dm.ptr, None)
        lib.bn_gs_color_zero_colour(vt.encode(), n, c, dcol.ptr, ddn.ptr, db.ptr, dcl.ptr, None)
        gm, gc = dm.numpy(), dcl.numpy()
        rows = perm[lo:]  # the rows the colour-order form writes
        assert np.array_equal(gc[rows].view(bits), gm[rows].view(bits))
        assert np.array_equal(gc[colour != c].view(bits), np.zeros(int(np.sum(colour != c)), dt).view(bits))


def colour_blocks(offd, col_ptr):
    """Row blocks of a colour-ordered copy, colour by colour (the solver's
    construction): boundaries and entry offsets, first block per colour."""
    bk, cb = [0], []
    for c in range(len(col_ptr) - 1):
        lo, hi = int(col_ptr[c]), int(col_ptr[c + 1])
        cb.append(len(bk) - 1)
        if hi > lo:
            sub = np.ascontiguousarray(offd.indptr[lo:hi + 1], np.int32)
            out = np.zeros(hi - lo + 1, np.int32)
            k = lib.bn_row_blocks(ip(sub), hi - lo, 2048, 256, ip(out))
            bk.extend(int(v) + lo for v in out[1:k + 1])
    cb.append(len(bk) - 1)
    bk = np.array(bk, np.int32)
    return bk, offd.indptr[bk].astype(np.int32), cb


@pytest.mark.parametrize("vt", ["d", "s"])
@pytest.mark.parametrize("which", ["network", "hubs"])
def test_row_block_colour_kernel_is_exact_gauss_seidel(vt, which):
    # the row-block colour kernel on the solver's layout: rows in colour
    # order holding only their off-diagonal entries, vectors natural, row r
    # updating map[r]; a forward sweep solves (D + L_pi) x_new = b - U_pi x_old
    if which == "network":
        net = make_network(14, 8, seed=3)
        A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    else:
        ia, ja, val = hub_matrix(3)
        H = sp.csr_matrix((val, ja, ia), shape=(len(ia) - 1, len(ia) - 1))
        A = (H + H.T).tocsr()
        A.setdiag(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
        A = A.tocsr()
        A.sort_indices()
    n = A.shape[0]
    aia, aja, ava = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    m = L.CSR(n, n, len(aja), ip(aia), ip(aja), dp(ava))
    perm = np.zeros(n, np.int32)
    cptr = C.POINTER(C.c_int)()
    mis = (np.arange(n) % 3 == 0).astype(np.int32)
    ncol = lib.bn_colour_order(C.byref(m), ip(mis), 1, ip(perm), C.byref(cptr))
    col_ptr = np.ctypeslib.as_array(cptr, (ncol + 1,)).copy()
    Ac = A[perm].tocsr()
    Ac.sort_indices()
    coo = Ac.tocoo()
    keep = coo.col != perm[coo.row]                     # drop each row's own diagonal
    offd = sp.csr_matrix((coo.data[keep], (coo.row[keep], coo.col[keep])), shape=(n, n))
    offd.sort_indices()
    bk, bq, cb = colour_blocks(offd, col_ptr)
    lens = np.diff(offd.indptr)
    case(f"row-block colour GS {which} n={n} colours={ncol} blocks={len(bk) - 1} "
         f"longest row {lens.max()} vt={vt}")
    rng = np.random.default_rng(4)
    b = rng.random(n) * 1e-3
    x_old = rng.random(n) * 1e3
    zbase = rng.random(n)
    D = A.diagonal()
    dt = np.float64 if vt == "d" else np.float32
    dia, dja, dv = dev(offd.indptr.astype(np.int32)), dev(offd.indices.astype(np.int32)), dev(offd.data.astype(dt))
    dbk, dbq, dmap, ddinv = dev(bk), dev(bq), dev(perm), dev((1.0 / D[perm]).astype(dt))
    db, dx, dx2 = dev(b.astype(dt)), dev(x_old.astype(dt)), dev(x_old.astype(dt))
    dzb, dzs = dev(zbase.astype(dt)), dev(np.zeros(n, dt))
    for c in range(ncol):
        lib.bn_gs_color_blocks(vt.encode(), cb[c], cb[c + 1] - cb[c], dbk.ptr, dbq.ptr, dia.ptr, dja.ptr, dv.ptr,
                               dmap.ptr, ddinv.ptr, db.ptr, dx.ptr, dzb.ptr, dzs.ptr, None)
        lib.bn_gs_color_map(vt.encode(), int(col_ptr[c]), int(col_ptr[c + 1]), dmap.ptr, dia.ptr, dja.ptr, dv.ptr,
                            ddinv.ptr, db.ptr, dx2.ptr, 4, None, None, None)
    x_new = dx.numpy().astype(np.float64)
    pos = np.empty(n, np.int64)
    pos[perm] = np.arange(n)
    ac = A.tocoo()
    lower = pos[ac.col] < pos[ac.row]
    upper = pos[ac.col] > pos[ac.row]
    Lp = sp.csr_matrix((ac.data[lower], (ac.row[lower], ac.col[lower])), shape=(n, n))
    Up = sp.csr_matrix((ac.data[upper], (ac.row[upper], ac.col[upper])), shape=(n, n))
    tol = 1e-12 if vt == "d" else 2e-5
    assert relerr(D * x_new + Lp @ x_new, b - Up @ x_old) < tol
    assert relerr(dzs.numpy().astype(np.float64), zbase + x_new) < (1e-15 if vt == "d" else 1e-6)
    # the lanes kernel on the same diagonal-free copy gives the same sweep
    assert relerr(dx2.numpy().astype(np.float64), x_new) < (1e-12 if vt == "d" else 1e-5)


@pytest.mark.parametrize("vt", ["d", "s"])
@pytest.mark.parametrize("backward", [0, 1])
@pytest.mark.parametrize("zero", [False, True])
@pytest.mark.parametrize("lanes", [1, 4])
def test_block_gs_sweep_solves_its_splitting(vt, backward, zero, lanes):
    # One block sweep over chunked rows is a residual correction whose
    # splitting matrix is the chunk-local lower (upper when backward)
    # triangle plus the diagonal with the l1 sum of the off-chunk couplings:
    #   (D1 + L_in) (x_new - x_old) = b - A x_old
    from test_schedule import chunk_order
    net = make_network(14, 8, seed=3)
    n = net.n
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(n, n))
    mis = (np.arange(n) % 3 == 0).astype(np.int32)
    nk, perm, chunk, ccol, cbnd = chunk_order(net.ia, net.ja, net.val, mis, 40)
    Ap = A[perm][:, perm].tocsr()
    Ap.sort_indices()
    chunk_of = np.repeat(np.arange(nk), np.diff(chunk))
    coo = Ap.tocoo()
    inside = chunk_of[coo.row] == chunk_of[coo.col]
    D = Ap.diagonal()
    offsum = np.zeros(n)
    np.add.at(offsum, coo.row[~inside], np.abs(coo.data[~inside]))
    D1 = D + offsum
    case(f"block GS n={n} chunks={nk} vt={vt} backward={backward} zero={zero} lanes={lanes}")
    rng = np.random.default_rng(4)
    b = rng.random(n) * 1e-9
    x_old = np.zeros(n) if zero else rng.random(n) * 1e3
    zbase = rng.random(n) * 1e3
    dt = np.float64 if vt == "d" else np.float32
    dia, dja, dv = dev(Ap.indptr.astype(np.int32)), dev(Ap.indices.astype(np.int32)), dev(Ap.data.astype(dt))
    dbd, db = dev((1.0 / D1).astype(dt)), dev(b.astype(dt))
    dxo, dxn = dev(x_old.astype(dt)), dev(np.zeros(n, dt))
    dzb, dzs = dev(zbase.astype(dt)), dev(np.zeros(n, dt))
    dch, dcc, dcb = dev(chunk), dev(ccol), dev(cbnd)
    lib.bn_block_gs(vt.encode(), nk, int(np.diff(chunk).max()), dch.ptr, dcc.ptr, dcb.ptr, dia.ptr, dja.ptr, dv.ptr,
                    dbd.ptr, db.ptr, None if zero else dxo.ptr, dxn.ptr, backward, lanes, dzb.ptr, dzs.ptr, None)
    x_new = dxn.numpy().astype(np.float64)
    tri = sp.triu(Ap, k=1) if backward else sp.tril(Ap, k=-1)
    tri = tri.tocoo()
    keep = chunk_of[tri.row] == chunk_of[tri.col]
    T_in = sp.csr_matrix((tri.data[keep], (tri.row[keep], tri.col[keep])), shape=(n, n))
    e = x_new - x_old
    lhs = D1 * e + T_in @ e
    rhs = b - Ap @ x_old
    assert relerr(lhs, rhs) < (1e-12 if vt == "d" else 2e-5)
    assert relerr(dzs.numpy().astype(np.float64), zbase + x_new) < (1e-15 if vt == "d" else 1e-6)


@pytest.mark.parametrize("vt", ["d", "s"])
@pytest.mark.parametrize("lanes", [1, 4])
def test_scaled_jacobi_sweep_writes_next_first_step(vt, lanes):
    # The scaled sweep writes the plain sweep's x bit for bit and also
    # x_i dinv_i, which is what the diagonal scaling of the next triangle's
    # first step forms from that x
    net = make_network(14, 8, seed=5)
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    Lo = sp.tril(A).tocsr()
    Lo.sort_indices()
    n = net.n
    case(f"scaled Jacobi sweep n={n} vt={vt} lanes={lanes}")
    dt = np.float64 if vt == "d" else np.float32
    rng = np.random.default_rng(7)
    b, xin = rng.random(n), rng.random(n)
    dinv = (1.0 / Lo.diagonal()).astype(dt)
    dia, dja, dv = dev(Lo.indptr.astype(np.int32)), dev(Lo.indices.astype(np.int32)), dev(Lo.data.astype(dt))
    dd, db, dxi = dev(dinv), dev(b.astype(dt)), dev(xin.astype(dt))
    x1, x2, xs = dev(np.zeros(n, dt)), dev(np.zeros(n, dt)), dev(np.zeros(n, dt))
    lib.bn_jacobi_rows(vt.encode(), n, dia.ptr, dja.ptr, dv.ptr, dd.ptr, db.ptr, dxi.ptr, x1.ptr, lanes, None, None,
                       None)
    lib.bn_jacobi_rows_scaled(vt.encode(), n, dia.ptr, dja.ptr, dv.ptr, dd.ptr, db.ptr, dxi.ptr, x2.ptr, xs.ptr,
                              lanes, None)
    bits = np.uint64 if vt == "d" else np.uint32
    g1, g2, gs = x1.numpy(), x2.numpy(), xs.numpy()
    assert np.array_equal(g1.view(bits), g2.view(bits))
    assert np.array_equal(gs.view(bits), (g2 * dinv).astype(dt).view(bits))
    # and the sweep itself: x = D^-1 (b - L_off xin)
    off = (Lo - sp.diags(Lo.diagonal())).tocsr()
    ref = (b - off @ xin) / Lo.diagonal()
    assert relerr(g1.astype(np.float64), ref) < (1e-14 if vt == "d" else 1e-6)


def test_level_scheduled_ic_against_oracle(oracle):
    net = make_network(16, 9)
    h = oracle.build(net.ia, net.ja, net.val, coarsening="rs")
    lia, lja, lv = h.fine_ichol()
    n = net.n
    rng = np.random.default_rng(5)
    b = rng.random(n)
    ref = np.zeros(n)
    oracle.lib.h_ichol_solve(h.h, 0, dp(b), dp(ref))
    Lm = sp.csr_matrix((lv, lja, lia), shape=(n, n))
    U = Lm.T.tocsr()
    U.sort_indices()
    dinv = 1.0 / Lm.diagonal()
    rows_f = np.zeros(n, np.int32)
    rows_b = np.zeros(n, np.int32)
    pf, pb = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    nf = lib.bn_level_sets(ip(lia), ip(lja), n, 1, ip(rows_f), C.byref(pf))
    ui, uj = U.indptr.astype(np.int32), U.indices.astype(np.int32)
    nb = lib.bn_level_sets(ip(ui), ip(uj), n, 0, ip(rows_b), C.byref(pb))
    ptr_f = np.ctypeslib.as_array(pf, (nf + 1,)).copy()
    ptr_b = np.ctypeslib.as_array(pb, (nb + 1,)).copy()
    case(f"level-scheduled IC n={n} levels={nf}/{nb}")
    d_lia, d_lja, d_lv = dev(lia), dev(lja), dev(lv)
    d_uia, d_uja, d_uv = dev(ui), dev(uj), dev(U.data)
    d_rf, d_rb, d_dinv, d_b = dev(rows_f), dev(rows_b), dev(dinv), dev(b)
    d_y, d_x = dev(np.zeros(n)), dev(np.zeros(n))
    for k in range(nf):
        lib.bn_trsv_level(b"d", int(ptr_f[k + 1] - ptr_f[k]), d_rf.ptr + 4 * int(ptr_f[k]), d_lia.ptr,
                          d_lja.ptr, d_lv.ptr, d_dinv.ptr, d_b.ptr, d_y.ptr, 2, None, None, None)
    for k in range(nb):
        lib.bn_trsv_level(b"d", int(ptr_b[k + 1] - ptr_b[k]), d_rb.ptr + 4 * int(ptr_b[k]), d_uia.ptr,
                          d_uja.ptr, d_uv.ptr, d_dinv.ptr, d_y.ptr, d_x.ptr, 2, None, None, None)
    assert relerr(d_x.numpy(), ref) < 1e-12


def test_dense_gemv():
    rng = np.random.default_rng(6)
    for n in (1, 7, 300, 1051):
        case(f"dense gemv n={n}")
        M = rng.random((n, n))
        x = rng.random(n)
        dM, dx, dy = dev(M), dev(x), dev(np.zeros(n))
        lib.bn_dense_gemv(b"d", n, dM.ptr, dx.ptr, dy.ptr, None)
        assert relerr(dy.numpy(), M @ x) < 1e-13


@pytest.mark.parametrize("block", [False, True])
def test_preconditioner_symmetric_and_matches_library_entry(net30, block):
    case(f"preconditioner symmetry block={block}")
    s = AMGSolver(net30.ia, net30.ja, net30.val, min_size=500, block_smoother=block, chunk_rows=512)
    rng = np.random.default_rng(7)
    u, v = rng.random(net30.n), rng.random(net30.n)
    Bu, Bv = s.precondition(u), s.precondition(v)
    a, b = v @ Bu, u @ Bv
    assert abs(a - b) <= 1e-10 * abs(a)
    # the plain V-cycle as well (amg_vcycle_precond)
    Vu, Vv = s.precondition(u, "vcycle"), s.precondition(v, "vcycle")
    assert abs(v @ Vu - u @ Vv) <= 1e-10 * abs(v @ Vu)
