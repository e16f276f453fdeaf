"""The reference's own driver (C_amg1/test_amg.c, unchanged) as a caller:
built once against the reference sources and once against include/amg.h
and libbnamg.so (oracle/_ref, see oracle/Makefile), run on the same COO
matrix file and rhs.txt.  With libbnamg its PCG runs on the GPU (SpMV and
BLAS-1 in amg_pcg, the AMG + IC(0) preconditioner through the cached
device mirror of its amg_data); both must converge, with iteration counts
within a few of each other."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from blood_network_amd import make_network

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
DRIVERS = {k: os.path.join(REF, f"test_amg_{k}") for k in ("ref", "bn")}


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def write_coo(path, ia, ja, val):
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    with open(path, "w") as f:
        f.write(f"{n} {n} {len(ja)}\n")
        np.savetxt(f, np.column_stack([rows + 1, ja + 1, val]), fmt="%d %d %.17g")


def run(kind, cwd, args):
    r = subprocess.run([DRIVERS[kind]] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    m = re.search(r"iterations = (-?\d+), relres = ([0-9.eE+-]+)", r.stderr)
    return r, (int(m.group(1)), float(m.group(2))) if m else None


@pytest.mark.parametrize("two_matrix", [False, True])
def test_reference_driver_on_libbnamg(tmp_path, two_matrix):
    if not all(os.path.exists(p) for p in DRIVERS.values()):
        pytest.skip("drop-in drivers not built (oracle/Makefile dropin)")
    net = make_network(24, 10, seed=3)
    case(f"C_amg1 test_amg driver, n={net.n}, two-matrix={two_matrix}")
    write_coo(str(tmp_path / "A.txt"), net.ia, net.ja, net.val)
    with open(tmp_path / "rhs.txt", "w") as f:
        f.write(f"{net.n}\n")
        np.savetxt(f, net.b, fmt="%.17g")
    args = ["A.txt"]
    if two_matrix:
        # hierarchy and IC(0) from a matrix with other radii
        rng = np.random.default_rng(5)
        vp, _ = net.assemble(net.conductance(net.radius * np.exp(0.2 * rng.normal(size=net.n_edges))))
        write_coo(str(tmp_path / "P.txt"), net.ia, net.ja, vp)
        args.append("P.txt")
    args += ["0.25", "1", "4096"]
    res = {}
    for kind in ("ref", "bn"):
        r, got = run(kind, str(tmp_path), args)
        print(f"--- {kind}: rc={r.returncode}\n{r.stderr[-1500:]}", flush=True)
        assert r.returncode == 0 and got is not None, kind
        assert "RHS loaded from rhs.txt" in r.stderr
        res[kind] = got
    (it_r, rr_r), (it_b, rr_b) = res["ref"], res["bn"]
    print(f"iterations reference={it_r} libbnamg={it_b}; relres {rr_r:.2e} / {rr_b:.2e}")
    assert it_b > 0 and rr_b < 1e-5
    assert abs(it_b - it_r) <= 3


def test_network_case_files_to_solution_files(tmp_path):
    """A Cartesian case written as <base>.fMx/.pMx/.dia/.BC, read and
    assembled by libbnamg (read_nwk.m semantics), solved on the GPU, and
    written as p_interior / labelled files (test_amg.m:68-72,
    label_solution.m:42-47); the pressures match a direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from blood_network_amd import AMGSolver, read_network_case, write_labelled, write_pressures
    from test_nwk_io import cartesian_case, write_case

    faces, pts, bc = cartesian_case(40, 30, 20, h=(2e-5, 2e-5, 4e-5), seed=2)
    base = str(tmp_path / "nwk_Cartesian_40x30x20")
    write_case(base, faces, pts, bc, dia=np.full(len(faces), 5e-6))
    c = read_network_case(base)
    case(f"network case files: {c.n_points} points, {c.n_faces} faces, n={c.n}")
    s = AMGSolver(c.ia, c.ja, c.val, 0.25, 128, 20, 2)
    p, it = s.solve(c.b, tol=1e-10, maxiter=500)
    A = sp.csr_matrix((c.val, c.ja, c.ia), shape=(c.n, c.n))
    ref = spla.spsolve(A.tocsc(), c.b)
    err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
    print(f"iterations={it} rel err vs direct solve {err:.2e}")
    assert it > 0 and err < 1e-8
    write_pressures(tmp_path / "p_interior_40x30x20.txt", p)
    write_labelled(tmp_path / "p_interior_40x30x20_labeled.txt", c.interior, p)
    assert np.array_equal(np.loadtxt(tmp_path / "p_interior_40x30x20.txt"), p)
    lab = np.loadtxt(tmp_path / "p_interior_40x30x20_labeled.txt")
    assert np.array_equal(lab[:, 0].astype(int), c.interior) and np.array_equal(lab[:, 1], p)
