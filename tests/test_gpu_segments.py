"""Many uneven networks in one solve: the block-diagonal union with one
hierarchy and per-network PCG scalars.  Every network's pressures against
its own sparse direct solve."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from blood_network_amd import AMGSolver, uneven_batch, union_networks

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def batch():
    nets = uneven_batch(12, seed=5, edge=(6, 16), depth=(4, 8))
    return nets, union_networks(nets)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_network_matches_its_direct_solve(batch, precision):
    nets, (ia, ja, val, b, seg) = batch
    case(f"{len(nets)} networks, n={seg[-1]}, {precision}")
    s = AMGSolver(ia, ja, val, precision=precision, fine_ic="jacobi", min_size=500)
    b = b.copy()
    b[seg[3]:seg[4]] = 0.0                      # a network with no driving pressure stops at once
    x, its = s.solve_segments(b, seg, tol=1e-12)
    print(f"iterations {its.tolist()}, union solve {s.report.solve_ms:.2f} ms")
    assert its[3] == 0 and not np.any(x[seg[3]:seg[4]])
    for g, net in enumerate(nets):
        if g == 3:
            continue
        A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
        xd = spla.spsolve(A.tocsc(), b[seg[g]:seg[g + 1]])
        assert 0 < its[g] < 200
        assert rel(x[seg[g]:seg[g + 1]], xd) < 1e-8, g


def test_graph_replay_and_reuse(batch, monkeypatch):
    nets, (ia, ja, val, b, seg) = batch
    case("segmented solve: graph vs eager, repeated")
    s = AMGSolver(ia, ja, val, precision="fp32", fine_ic="jacobi", min_size=500)
    monkeypatch.setenv("BN_NO_GRAPH", "1")
    xe, ie = s.solve_segments(b, seg, tol=1e-9)
    monkeypatch.delenv("BN_NO_GRAPH")
    for _ in range(2):
        xg, ig = s.solve_segments(b, seg, tol=1e-9)
        assert np.array_equal(ig, ie) and np.array_equal(xg, xe)
    # the plain solve of the union is unaffected by the cached segmentation
    xu, iu = s.solve(b, tol=1e-9)
    assert iu > 0 and rel(xu, xe) < 1e-6


def test_bad_segmentation_is_rejected(batch):
    nets, (ia, ja, val, b, seg) = batch
    s = AMGSolver(ia, ja, val, fine_ic="jacobi", min_size=500)
    bad = seg.copy()
    bad[-1] -= 1
    with pytest.raises(RuntimeError, match="seg_ptr"):
        s.solve_segments(b, bad)
    bad = seg.copy()
    bad[2], bad[3] = bad[3], bad[2]
    with pytest.raises(RuntimeError, match="non-decreasing"):
        s.solve_segments(b, bad)
