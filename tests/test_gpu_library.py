"""The library-level entry points on the reference's own host structures
(include/amg.h: amg_build_hierarchy, amg_backslash, amg_fwdslash,
amg_vcycle_precond, amg_ichol_precond, ichol_solve, amg_pcg), as a C
caller of libamg uses them: the hierarchy is built on the host, mirrored
on the GPU on first use and the work runs there."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from blood_network_amd import AMGSolver, make_network
from blood_network_amd import _lib as L
from oracle_lib import dp, ip

pytestmark = pytest.mark.gpu
lib = L.lib
PRECOND = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))
lib.amg_pcg.restype = C.c_int
lib.amg_pcg.argtypes = [C.POINTER(L.CSR), L.DP, L.DP, PRECOND, C.c_void_p, C.c_double, C.c_int, C.c_int]


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class Hierarchy:
    """amg_data + IC(0) built by the library's host set-up, options as the
    reference wrapper sets them (C_amg1/src/amg_wrapper.c:80-89)."""

    def __init__(self, net, min_size=500, nu=1):
        self.ia, self.ja, self.val = net.ia, net.ja, net.val
        self.n = net.n
        self.A = L.CSR(net.n, net.n, net.nnz, ip(self.ia), ip(self.ja), dp(self.val))
        self.amg = L.AMGData()
        self.amg.threshold = 0.25
        self.amg.min_size = min_size
        self.amg.max_levels = 20
        self.amg.nu = nu
        self.amg.smoother = L.SMOOTHER_CF_GS
        self.amg.omega_jacobi = 0.5
        self.amg.coarsening = L.COARSEN_PMIS
        lib.amg_build_hierarchy(C.addressof(self.amg), C.byref(self.A))
        self.L = L.IcholFactor()
        lib.ichol_compute(C.byref(self.A), C.addressof(self.L))

    @property
    def ptr(self):
        return C.addressof(self.amg)

    def level_matrix(self, lev):
        m = self.amg.levels[lev].A
        ia = np.ctypeslib.as_array(m.ia, (m.row + 1,)).copy()
        ja = np.ctypeslib.as_array(m.ja, (m.nnz,)).copy()
        va = np.ctypeslib.as_array(m.val, (m.nnz,)).copy()
        return sp.csr_matrix((va, ja, ia), shape=(m.row, m.row))

    def free(self):
        lib.amg_free(self.ptr)
        lib.ichol_free(C.addressof(self.L))


@pytest.fixture(scope="module")
def hier(net30):
    h = Hierarchy(net30)
    yield h
    h.free()


def test_coarsest_half_cycles_solve_exactly(hier):
    nl = hier.amg.num_levels
    assert nl >= 3
    lev = nl - 1
    Ac = hier.level_matrix(lev)
    rng = np.random.default_rng(1)
    b = rng.random(Ac.shape[0])
    ref = spla.spsolve(Ac.tocsc(), b)
    for name, fn in (("backslash", lib.amg_backslash), ("fwdslash", lib.amg_fwdslash)):
        case(f"amg_{name} on the coarsest level n={Ac.shape[0]}")
        x = np.zeros(Ac.shape[0])
        fn(hier.ptr, lev, dp(b), dp(x), 1)
        assert rel(x, ref) < 1e-10


def test_vcycle_is_backslash_then_fwdslash(hier):
    # amg_vcycle_precond: u = B\\g from zero, then u + F\\(g - A u)
    case("amg_vcycle_precond vs amg_backslash + amg_fwdslash")
    A = hier.level_matrix(0)
    rng = np.random.default_rng(2)
    g = rng.random(hier.n)
    z = np.zeros(hier.n)
    lib.amg_vcycle_precond(hier.ptr, dp(g), dp(z))
    u = np.zeros(hier.n)
    lib.amg_backslash(hier.ptr, 0, dp(g), dp(u), 1)
    r = g - A @ u
    c = np.zeros(hier.n)
    lib.amg_fwdslash(hier.ptr, 0, dp(r), dp(c), 1)
    assert rel(u + c, z) < 1e-12
    # a convergent symmetric V-cycle contracts the error in the energy norm:
    # e = (I - B A) x for one application from zero
    x_true = spla.spsolve(A.tocsc(), g)
    e = x_true - z
    assert e @ (A @ e) < 0.81 * (x_true @ (A @ x_true))


def test_library_preconditioners_match_the_handle(hier, net30):
    case("library-level preconditioners vs the handle API")
    # same input, no symmetrisation: the handle then builds the identical hierarchy
    s = AMGSolver(net30.ia, net30.ja, net30.val, 0.25, 500, 20, 1, symmetrize=False)
    rng = np.random.default_rng(3)
    g = rng.random(net30.n)
    z = np.zeros(net30.n)
    lib.amg_vcycle_precond(hier.ptr, dp(g), dp(z))
    assert rel(z, s.precondition(g, "vcycle")) < 1e-12
    lib.amg_ichol_precond(hier.ptr, C.byref(hier.A), C.addressof(hier.L), dp(g), dp(z))
    assert rel(z, s.precondition(g)) < 1e-12


def test_ichol_solve_matches_oracle(hier, net30, oracle):
    case("ichol_solve (library entry) vs the oracle")
    h = oracle.build(net30.ia, net30.ja, net30.val, coarsening="pmis", min_size=500, symmetrize=False)
    rng = np.random.default_rng(4)
    b = rng.random(net30.n)
    ref = np.zeros(net30.n)
    oracle.lib.h_ichol_solve(h.h, 0, dp(b), dp(ref))
    x = np.zeros(net30.n)
    lib.ichol_solve(C.addressof(hier.L), dp(b), dp(x))
    assert rel(x, ref) < 1e-12


def test_amg_pcg_with_host_preconditioner(hier, net30):
    # amg_pcg runs SpMV and BLAS-1 on the device and calls the caller's
    # preconditioner on host vectors each step (C_amg1/src/amg_pcg.c)
    case("amg_pcg with amg_ichol_precond as the host callback")

    @PRECOND
    def precond(data, gp, zp):
        lib.amg_ichol_precond(hier.ptr, C.byref(hier.A), C.addressof(hier.L), C.cast(gp, L.DP),
                              C.cast(zp, L.DP))

    x = np.zeros(net30.n)
    it = lib.amg_pcg(C.byref(hier.A), dp(net30.b), dp(x), precond, None, 1e-10, 500, 0)
    xd = spla.spsolve(hier.level_matrix(0).tocsc(), net30.b)
    s = AMGSolver(net30.ia, net30.ja, net30.val, 0.25, 500, 20, 1, symmetrize=False)
    xh, ih = s.solve(net30.b, tol=1e-10, maxiter=500)
    print(f"amg_pcg iterations {it}, handle {ih}, rel(x, direct) {rel(x, xd):.2e}")
    assert it > 0 and abs(it - ih) <= 1
    assert rel(x, xd) < 1e-8


def test_amg_pcg_reads_the_matrix_afresh():
    # one dcsr_mat struct, solved three times: as built, with its values
    # changed in place, then pointing at another (larger) matrix; the
    # reference reads A on every call, so must the device copy
    case("amg_pcg on one dcsr_mat: new values in place, then a new size")
    jac = {}

    @PRECOND
    def precond(data, gp, zp):
        d = jac["d"]
        g = np.ctypeslib.as_array(gp, (len(d),))
        z = np.ctypeslib.as_array(zp, (len(d),))
        z[:] = g / d

    def solve_and_check(A, m, b):
        jac["d"] = m.diagonal()
        x = np.zeros(m.shape[0])
        it = lib.amg_pcg(C.byref(A), dp(b), dp(x), precond, None, 1e-11, 20000, 0)
        xd = spla.spsolve(m.tocsc(), b)
        print(f"n={m.shape[0]} iterations {it} rel(x, direct) {rel(x, xd):.2e}")
        assert it > 0 and rel(x, xd) < 1e-7

    n1 = make_network(10, 6, seed=3)
    ia, ja, val = n1.ia.copy(), n1.ja.copy(), n1.val.copy()
    A = L.CSR(n1.n, n1.n, n1.nnz, ip(ia), ip(ja), dp(val))
    solve_and_check(A, sp.csr_matrix((val, ja, ia), shape=(n1.n, n1.n)), n1.b)
    val *= 2.0                                   # in place: same pointers
    solve_and_check(A, sp.csr_matrix((val, ja, ia), shape=(n1.n, n1.n)), n1.b)
    n2 = make_network(12, 6, seed=4)
    ia2, ja2, val2 = n2.ia.copy(), n2.ja.copy(), n2.val.copy()
    A.row = A.col = n2.n
    A.nnz = n2.nnz
    A.ia, A.ja, A.val = ip(ia2), ip(ja2), dp(val2)
    solve_and_check(A, sp.csr_matrix((val2, ja2, ia2), shape=(n2.n, n2.n)), n2.b)


def test_two_handles_on_two_threads():
    # handles are independent (one HIP stream each, no shared state); a
    # handle's results do not depend on another thread working beside it
    import threading
    case("two solver handles solved concurrently on two threads")
    na, nb = make_network(16, 9, seed=1), make_network(14, 8, seed=2)
    sa = AMGSolver(na.ia, na.ja, na.val, coarsening="pmis", min_size=500)
    sb = AMGSolver(nb.ia, nb.ja, nb.val, coarsening="hmis", min_size=500)
    ref_a, _ = sa.solve(na.b, tol=1e-10)
    ref_b, _ = sb.solve(nb.b, tol=1e-10)
    out, errs = {}, []

    def work(key, s, b):
        try:
            for rep in range(4):
                out[(key, rep)] = s.solve(b, tol=1e-10)[0]
        except Exception as e:  # reported below
            errs.append(e)

    ts = [threading.Thread(target=work, args=("a", sa, na.b)), threading.Thread(target=work, args=("b", sb, nb.b))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for rep in range(4):
        assert np.array_equal(out[("a", rep)], ref_a)
        assert np.array_equal(out[("b", rep)], ref_b)
    sa.free()
    sb.free()


def test_library_preconditioner_from_two_threads(hier):
    # two threads on one amg_data: the mirror's lock runs them one after the
    # other, each gets the single-thread result
    import threading
    case("amg_vcycle_precond on one hierarchy from two threads")
    rng = np.random.default_rng(9)
    gs = [rng.random(hier.n) for _ in range(2)]
    ref = []
    for g in gs:
        z = np.zeros(hier.n)
        lib.amg_vcycle_precond(hier.ptr, dp(g), dp(z))
        ref.append(z)
    got = [[None] * 3 for _ in range(2)]

    def work(k):
        for rep in range(3):
            z = np.zeros(hier.n)
            lib.amg_vcycle_precond(hier.ptr, dp(gs[k]), dp(z))
            got[k][rep] = z

    ts = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for k in range(2):
        for rep in range(3):
            assert np.array_equal(got[k][rep], ref[k])


def test_set_values_device_reaches_the_fp32_preconditioner(net16):
    # the residuals inside the FP32 AMG+IC preconditioner use the solve
    # matrix (amg_ichol_precond.c:37,44): after new values on the device they
    # must use the new ones, as the FP64 preconditioner does
    case("set_values_device: FP32 and FP64 preconditioners see the same new matrix")
    net = net16
    s = AMGSolver(net.ia, net.ja, net.val, coarsening="pmis", min_size=500, precision="fp32", fine_ic="jacobi")
    g = np.random.default_rng(10).random(net.n)
    before = s.precondition(g)
    dv = L.DeviceArray(net.val * 1.7)
    s.set_values_device(net.nnz, dv.ptr)
    z32 = s.precondition(g)
    s.set_mode(precision="fp64", fine_ic="jacobi")
    z64 = s.precondition(g)
    e, moved = rel(z32, z64), rel(z32, before)
    print(f"FP32 vs FP64 after new values {e:.2e}; change from before {moved:.2e}")
    assert e < 1e-4 and moved > 1e-2
    dv.free()
    s.free()
