"""Vessel-network case files and solution files (include/bn_netgen.h
bn_nwk_*, bn_write_*): the interior system against a scipy restatement of
Matlab_amg/read_nwk.m:15-37 (with ConnectivityMx.m and
makeDiffusiveConductance.m), pinned by closed forms on Cartesian grids,
and the error paths.  The reference ships no case files (its
four_large_Cartesian_networks directory is absent), so cases are written
here."""
import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import read_network_case, write_labelled, write_pressures


def read_nwk_reference(faces, pts, bc):
    """read_nwk.m restated with scipy: faces (nf, 2) and bc nodes 1-based."""
    nf, np_ = len(faces), len(pts)
    fi = np.arange(nf)
    C1 = sp.csr_matrix((np.r_[np.ones(nf), -np.ones(nf)], (np.r_[fi, fi], np.r_[faces[:, 0], faces[:, 1]] - 1)),
                       shape=(nf, np_))                                       # ConnectivityMx.m:1-5
    vec = pts[faces[:, 1] - 1] - pts[faces[:, 0] - 1]                           # makeDiffusiveConductance.m:4
    h = [np.linalg.norm(vec[np.nonzero(vec[:, d] > 0)[0][0]]) for d in range(3)]
    cond = np.full(nf, np.nan)
    cond[vec[:, 0] > 0] = h[1] * h[2] / h[0]
    cond[vec[:, 1] > 0] = h[0] * h[2] / h[1]
    cond[vec[:, 2] > 0] = h[0] * h[1] / h[2]
    W = sp.diags(cond)
    bnodes = bc[:, 0].astype(int)
    interior = np.setdiff1d(np.arange(1, np_ + 1), bnodes)                      # read_nwk.m:20-22
    CI, CB = C1[:, interior - 1], C1[:, bnodes - 1]
    pB = bc[:, 2]
    A = (CI.T @ W @ CI).tocsr()
    A.eliminate_zeros()
    A.sort_indices()
    b = -(CI.T @ W @ CB @ pB)                                                   # read_nwk.m:32-34
    return A, np.asarray(b).ravel(), interior


def cartesian_case(nx, ny, nz, h=(1.0, 1.0, 1.0), seed=0, shuffle=True):
    """Grid points, +x/+y/+z faces, Dirichlet pressures on the x = 0 and
    x = max planes (1-based ids, as the case files hold them)."""
    rng = np.random.default_rng(seed)
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.column_stack([xx.ravel() * h[0], yy.ravel() * h[1], zz.ravel() * h[2]])
    faces = [np.column_stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()]),
             np.column_stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()]),
             np.column_stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()])]
    faces = np.concatenate(faces) + 1
    if shuffle:
        faces = faces[rng.permutation(len(faces))]
    left, right = idx[:, :, 0].ravel() + 1, idx[:, :, -1].ravel() + 1
    bc = np.column_stack([np.r_[left, right], np.ones(len(left) + len(right)),
                          np.r_[np.full(len(left), 13300.0), np.full(len(right), 1330.0)]])
    return faces, pts, bc


def write_case(base, faces, pts, bc, dia=None):
    np.savetxt(f"{base}.fMx", np.column_stack([np.arange(1, len(faces) + 1), faces]), fmt="%d")
    np.savetxt(f"{base}.pMx", pts, fmt="%.17g")
    if dia is None:
        open(f"{base}.dia", "w").close()
    else:
        np.savetxt(f"{base}.dia", dia, fmt="%.17g")
    np.savetxt(f"{base}.BC", bc, fmt="%.17g")


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (2e-5, 5e-5, 3e-5)])
def test_cartesian_case_matches_read_nwk(tmp_path, h):
    faces, pts, bc = cartesian_case(7, 5, 4, h)
    base = str(tmp_path / "nwk")
    write_case(base, faces, pts, bc, dia=np.full(len(faces), 4e-6))
    c = read_network_case(base)
    A, b, interior = read_nwk_reference(faces, pts, bc)
    assert c.n == A.shape[0] == 5 * 5 * 4
    assert np.array_equal(c.interior, interior)
    assert np.array_equal(c.ia, A.indptr) and np.array_equal(c.ja, A.indices)
    assert np.allclose(c.val, A.data, rtol=1e-14, atol=0)
    assert np.allclose(c.b, b, rtol=1e-14, atol=1e-14 * np.abs(b).max())
    # closed form: x faces conduct hy hz / hx, so an interior row's diagonal
    # is the sum of its neighbours' face conductances and every row sums to
    # the conductance of its faces to Dirichlet nodes
    gx, gy, gz = h[1] * h[2] / h[0], h[0] * h[2] / h[1], h[0] * h[1] / h[2]
    Ad = sp.csr_matrix((c.val, c.ja, c.ia), shape=(c.n, c.n))
    off = -Ad.copy()
    off.setdiag(0)
    off.eliminate_zeros()
    assert set(np.round(np.unique(off.data) / min(gx, gy, gz), 9)) <= set(np.round(np.array([gx, gy, gz]) / min(gx, gy, gz), 9))
    rows = np.asarray(Ad.sum(axis=1)).ravel()
    xi = (c.interior - 1) % 7
    to_bc = np.where((xi == 1) | (xi == 5), gx, 0.0)
    assert np.allclose(rows, to_bc, rtol=1e-12, atol=1e-12 * gx)
    assert np.allclose(c.b[xi == 1], gx * 13300.0) and np.allclose(c.b[xi == 5], gx * 1330.0)
    assert c.diameters is not None and len(c.diameters) == len(faces)


def test_duplicate_faces_and_self_loops_sum_like_matlab(tmp_path):
    faces, pts, bc = cartesian_case(5, 4, 3)
    faces = np.concatenate([faces, faces[:17], [[8, 8]]])  # repeated faces, one self-loop
    base = str(tmp_path / "dup")
    write_case(base, faces, pts, bc)
    c = read_network_case(base)
    A, b, _ = read_nwk_reference(faces[:-1], pts, bc)  # the self-loop adds nothing
    assert np.array_equal(c.ia, A.indptr) and np.array_equal(c.ja, A.indices)
    assert np.allclose(c.val, A.data, rtol=1e-14) and np.allclose(c.b, b, rtol=1e-14)
    assert c.diameters is None


def test_case_errors(tmp_path):
    faces, pts, bc = cartesian_case(4, 3, 3, shuffle=False)
    base = str(tmp_path / "e")
    # a boundary node listed twice
    write_case(base, faces, pts, np.concatenate([bc, bc[:1]]))
    with pytest.raises(ValueError, match="twice"):
        read_network_case(base)
    # a face pointing in -x (MATLAB would give it a NaN conductance)
    f2 = faces.copy()
    f2[0] = f2[0][::-1]
    write_case(base, f2, pts, bc)
    with pytest.raises(ValueError, match="positive"):
        read_network_case(base)
    # a face naming a point that does not exist
    f3 = faces.copy()
    f3[3, 1] = len(pts) + 1
    write_case(base, f3, pts, bc)
    with pytest.raises(ValueError, match="names point"):
        read_network_case(base)
    # ragged rows
    write_case(base, faces, pts, bc)
    with open(f"{base}.pMx", "a") as f:
        f.write("1 2\n")
    with pytest.raises(ValueError, match="numbers"):
        read_network_case(base)
    # missing file
    with pytest.raises(ValueError, match="cannot open"):
        read_network_case(str(tmp_path / "absent"))


def test_solution_files_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    p = rng.normal(size=500) * 1e4
    nodes = np.sort(rng.choice(10**6, 500, replace=False)) + 1
    write_pressures(tmp_path / "p.txt", p)
    write_labelled(tmp_path / "pl.txt", nodes, p)
    got = np.loadtxt(tmp_path / "p.txt")
    assert np.array_equal(got, p)  # %.16e keeps 17 significant digits
    lines = open(tmp_path / "p.txt").read().splitlines()
    assert lines[0] == "%.16e" % p[0]
    lab = np.loadtxt(tmp_path / "pl.txt")
    assert np.array_equal(lab[:, 0].astype(int), nodes) and np.array_equal(lab[:, 1], p)
    assert open(tmp_path / "pl.txt").readline() == "%d %.16e\n" % (nodes[0], p[0])
