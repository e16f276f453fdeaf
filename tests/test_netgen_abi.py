"""Synthetic network generator and the C ABI surface (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from blood_network_amd import WORKLOADS, make_network
from blood_network_amd._lib import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_sizes_match_workload_definitions():
    # n = L^3 + 2 (2^d - 2); edges = 3 L^2 (L-1) + 2 ((2^d - 2) + 2^(d-1));
    # nnz = n + 2 (edges - edges touching the 2 roots)
    for L, d in [(16, 9), (12, 6), (20, 10)]:
        net = make_network(L, d)
        n = L ** 3 + 2 * (2 ** d - 2)
        m = 3 * L * L * (L - 1) + 2 * ((2 ** d - 2) + 2 ** (d - 1))
        assert (net.n, net.n_edges, net.nnz) == (n, m, n + 2 * (m - 4))
    net = make_network(16, 9)
    assert (net.n, net.nnz, net.n_edges) == (5116, 31212, 13052)


def test_system_is_the_eliminated_weighted_laplacian(net16):
    net = net16
    A = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
    assert abs(A - A.T).max() == 0
    for i in range(net.n):  # sorted rows, diagonal present
        cols = net.ja[net.ia[i]:net.ia[i + 1]]
        assert np.all(np.diff(cols) > 0) and i in cols
    # A = C_I^T W C_I, b = -C_I^T W C_B p_B (Matlab_amg/read_nwk.m:35-37)
    m = net.n_edges
    eu, ev, g = net.eu, net.ev, net.g
    CI = sp.lil_matrix((m, net.n))
    CB = sp.lil_matrix((m, net.n_dirichlet))
    for e in range(m):
        for node, sgn in ((eu[e], 1.0), (ev[e], -1.0)):
            if node >= 0:
                CI[e, node] = sgn
            else:
                CB[e, -1 - node] = sgn
    W = sp.diags(g)
    A2 = (CI.T @ W @ CI).tocsr()
    b2 = -(CI.T @ W @ CB) @ net.p_dirichlet
    assert abs(A - A2).max() <= 1e-14 * abs(A).max()
    assert np.allclose(net.b, b2, rtol=1e-13)
    assert np.count_nonzero(net.b) == 4  # the roots' two children each
    # Poiseuille conductance
    assert np.allclose(g, np.pi * net.radius ** 4 / (8 * 3.5e-3 * net.length), rtol=1e-14)


def test_generator_statistics_and_determinism():
    a = make_network(20, 10, seed=0)
    b = make_network(20, 10, seed=0)
    c = make_network(20, 10, seed=1)
    assert a.sha256() == b.sha256() != c.sha256()
    lat = 3 * 20 * 20 * 19
    logr = np.log(a.radius[:lat] / 4e-6)
    assert abs(logr.mean()) < 0.01 and abs(logr.std() - 0.2) < 0.01
    tr = a.radius[lat:]
    assert tr.max() <= 500e-6 * 2 ** (-1 / 3) * 1.1 + 1e-12
    assert np.allclose(a.length[lat:], 20 * tr)
    # leaf attachments go to distinct lattice nodes
    leaves = 2 ** 9
    att = []
    for t in range(2):
        start = lat + t * (2 ** 10 - 2 + leaves) + (2 ** 10 - 2)
        att.append(a.ev[start:start + leaves])
    att = np.concatenate(att)
    assert len(np.unique(att)) == len(att) and att.max() < 20 ** 3


def test_assemble_reproduces_generator(net16):
    val, b = net16.assemble(net16.g)
    assert np.array_equal(val, net16.val) and np.array_equal(b, net16.b)


def test_workload_sizes_declared():
    assert set(WORKLOADS) >= {"lattice16_trees9", "lattice100_trees15", "lattice160_trees17"}


def test_library_exports_every_declared_symbol():
    names = set()
    for h in ("amg.h", "bn_amg_ext.h", "bn_netgen.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        # header-only helpers (static inline, defined in the header) are not exports
        txt = re.sub(r"static inline[^{;]*\{.*?\n\}\n", "", txt, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", txt):
            nm = m.group(1)
            if nm not in ("if", "sizeof", "defined", "void", "int", "double", "char", "float"):
                names.add(nm)
    # drop the function-pointer parameter names
    names -= {"precond"}
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = sorted(names - exported)
    assert not missing, missing
    lib = C.CDLL(LIB_PATH)
    for nm in names:
        getattr(lib, nm)


def test_solve_without_device_fails_loudly(net16):
    from blood_network_amd import AMGSolver, device_available
    if device_available():
        pytest.skip("a GPU is present")
    s = AMGSolver(net16.ia, net16.ja, net16.val)
    with pytest.raises(RuntimeError, match="no HIP device"):
        s.solve(net16.b)


def test_ctypes_mirrors_match_the_c_headers(tmp_path):
    """The Python structs that mirror include/*.h (used for the library-level
    entry points and the options/info blocks) have the C layouts."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    from blood_network_amd import _lib as L
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "amg.h"\n#include "bn_amg_ext.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(dcsr_mat), sizeof(ichol_factor),'
        ' sizeof(amg_level), sizeof(amg_data), offsetof(amg_data, num_levels), offsetof(amg_data, coarsening),'
        ' sizeof(bn_options), offsetof(bn_options, chunk_rows), sizeof(bn_solve_info)); return 0;}\n')
    exe = tmp_path / "sz"
    inc = os.path.join(root, "include")
    subprocess.run(["gcc", "-I", inc, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(L.CSR), C.sizeof(L.IcholFactor), C.sizeof(L.AMGLevel), C.sizeof(L.AMGData),
            L.AMGData.num_levels.offset, L.AMGData.coarsening.offset, C.sizeof(L.Options),
            L.Options.chunk_rows.offset, C.sizeof(L.SolveInfo)]
    assert got == want


def test_union_of_networks_is_block_diagonal():
    from blood_network_amd import union_networks, uneven_batch
    nets = uneven_batch(5, seed=3, edge=(4, 9), depth=(3, 6))
    ia, ja, val, b, seg = union_networks(nets)
    A = sp.csr_matrix((val, ja, ia), shape=(seg[-1], seg[-1]))
    assert np.array_equal(np.diff(seg), [net.n for net in nets])
    for g, net in enumerate(nets):
        blk = A[seg[g]:seg[g + 1]]
        # every entry of network g's rows lies in its own diagonal block
        assert blk.indices.min() >= seg[g] and blk.indices.max() < seg[g + 1]
        Ag = sp.csr_matrix((net.val, net.ja, net.ia), shape=(net.n, net.n))
        assert (blk[:, seg[g]:seg[g + 1]] != Ag).nnz == 0
        assert np.array_equal(b[seg[g]:seg[g + 1]], net.b)
    # sizes follow the requested ranges; every network fits its leaves
    Ls = [round((net.n_nodes - 2 * (2 ** net.params.depth - 1)) ** (1 / 3)) for net in nets]
    assert all(4 <= Le <= 9 for Le in Ls)


def test_boundary_rooted_trees_perfuse_their_entry_region():
    # 4 + 4 trees of depth 6 on a 20^3 lattice, each rooted at a boundary-face
    # node: its 32 leaves attach to the nearest free lattice nodes
    from blood_network_amd import make_network
    Le, d, T = 20, 6, 4
    net = make_network(Le, d, T, seed=3, boundary_roots=1)
    tn, leaves = 2 ** d - 1, 2 ** (d - 1)
    assert net.n == Le ** 3 + 2 * T * (tn - 1)
    e = 3 * Le * Le * (Le - 1)
    targets = []
    for t in range(2 * T):
        e += tn - 1
        ids = net.ev[e:e + leaves]
        e += leaves
        c = np.stack([ids % Le, (ids // Le) % Le, ids // (Le * Le)], 1)
        # the first target is the root's boundary-face node
        assert np.any((c[0] == 0) | (c[0] == Le - 1))
        # breadth-first from it: every later target is at least as far (in
        # lattice steps, up to nodes an earlier tree took) and all lie close
        dist = np.abs(c - c[0]).sum(1)
        assert dist.max() <= 2 * leaves ** (1 / 3) + 4
        targets.append(ids)
    allt = np.concatenate(targets)
    assert len(np.unique(allt)) == len(allt)          # distinct across trees
    # the uniform form spreads the same leaves over the whole lattice
    flat = make_network(Le, d, T, seed=3)
    ids = flat.ev[3 * Le * Le * (Le - 1) + tn - 1:3 * Le * Le * (Le - 1) + tn - 1 + leaves]
    c = np.stack([ids % Le, (ids // Le) % Le, ids // (Le * Le)], 1)
    assert (c.max(0) - c.min(0)).min() > Le // 2
