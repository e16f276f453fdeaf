"""Every smoother of amg_solver_set_smoother on the GPU against the oracle
(C_amg1/src/amg_cycle.c:21-175): the V-cycle halves, the plain V-cycle,
the AMG + IC(0) preconditioner and whole PCG solves.

L1-Jacobi and damped Jacobi are order-free; SMOOTHER_GS (natural order) and
SMOOTHER_CF_GS with gs_exact (the reference's cf_order) run the sequential
sweep level by level over its dependency DAG, so every one of them is the
reference's smoother and the results agree with the oracle's to rounding.
The coarsest level runs the reference's IC-PCG to 1e-12, as the oracle
does.  The C-typed twin of the MEX smoother is checked against the
oracle's restatement of cf_gauss_seidel_mex.c and a dense closed form."""
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

SMOOTHERS = [("l1", L.SMOOTHER_L1, 0.5), ("jacobi", L.SMOOTHER_JACOBI, 0.5), ("jacobi", L.SMOOTHER_JACOBI, 0.8),
             ("gs", L.SMOOTHER_GS, 0.5), ("cf_gs", L.SMOOTHER_CF_GS, 0.5)]
IDS = ["l1", "jacobi0.5", "jacobi0.8", "gs", "cf_gs_exact"]


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def net():
    return make_network(20, 9, seed=1)


def oracle_hier(oracle, net, smoother, omega, nu=1):
    h = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=300, nu=nu, smoother=smoother)
    h.set_smoother(smoother, omega)
    return h


def gpu_solver(net, name, omega, nu=1, **kw):
    s = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, nu, coarsening="pmis", coarse="icpcg", smoother=name,
                  gs_exact=(name == "cf_gs"), **kw)
    s.set_smoother(name, omega)
    return s


@pytest.mark.parametrize("name,code,omega", SMOOTHERS, ids=IDS)
def test_preconditioners_match_the_oracle(net, oracle, name, code, omega):
    case(f"V-cycle and AMG+IC preconditioner, smoother {name} omega {omega}")
    h = oracle_hier(oracle, net, code, omega)
    s = gpu_solver(net, name, omega)
    assert s.num_levels() == h.num_levels() >= 3
    rng = np.random.default_rng(5)
    g = rng.random(net.n) - 0.5
    zv = s.precondition(g, "vcycle")
    zi = s.precondition(g)
    ev, ei = rel(zv, h.vcycle(g)), rel(zi, h.precond(g))
    print(f"rel err vs oracle: vcycle {ev:.2e}, AMG+IC {ei:.2e}")
    assert ev < 1e-10 and ei < 1e-10
    s.free()


@pytest.mark.parametrize("name,code,omega", SMOOTHERS, ids=IDS)
def test_solves_match_the_oracle(net, oracle, name, code, omega):
    case(f"PCG solve, smoother {name} omega {omega}")
    h = oracle_hier(oracle, net, code, omega)
    xo, ito, _ = h.solve(net.b, 1e-10, 500)
    s = gpu_solver(net, name, omega)
    x, it = s.solve(net.b, tol=1e-10, maxiter=500)
    e = rel(x, xo)
    print(f"iterations gpu {it} oracle {ito}; rel(x, oracle) {e:.2e}")
    assert it > 0 and abs(it - ito) <= 1
    assert e < 1e-8
    s.free()


def test_exact_cf_sweep_differs_from_the_multicolour_sweep(net):
    # the two CF-GS forms are different smoothers: the exact one must not
    # silently fall back to the multicolour sweep (or the reverse)
    case("exact vs multicolour CF-GS")
    g = np.random.default_rng(6).random(net.n)
    a = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, 1, coarsening="pmis", coarse="icpcg", gs_exact=True)
    b = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, 1, coarsening="pmis", coarse="icpcg", gs_exact=False)
    d = rel(a.precondition(g, "vcycle"), b.precondition(g, "vcycle"))
    print(f"exact vs multicolour V-cycle: {d:.2e}")
    assert d > 1e-6
    a.free()
    b.free()


def test_omega_change_between_solves_on_one_handle(net, oracle):
    # the captured step graph holds omega as a kernel argument: a change
    # must reach the next solve
    case("damped Jacobi omega 0.5 -> 0.8 on one handle (graph replay)")
    s = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, 1, coarsening="pmis", smoother="jacobi")
    s.set_smoother("jacobi", 0.5)
    x1, it1 = s.solve(net.b, tol=1e-10, maxiter=500)
    x1b, _ = s.solve(net.b, tol=1e-10, maxiter=500)      # replay of the captured graph
    s.set_smoother("jacobi", 0.8)
    x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)
    fresh = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, 1, coarsening="pmis", smoother="jacobi")
    fresh.set_smoother("jacobi", 0.8)
    x3, it3 = fresh.solve(net.b, tol=1e-10, maxiter=500)
    h = oracle_hier(oracle, net, L.SMOOTHER_JACOBI, 0.8)
    g = np.random.default_rng(8).random(net.n)
    print(f"iterations: w=0.5 {it1}, then w=0.8 {it2}, fresh w=0.8 {it3}")
    assert np.array_equal(x1, x1b)
    assert it2 == it3 and rel(x2, x3) < 1e-10
    assert rel(s.precondition(g, "vcycle"), h.vcycle(g)) < 1e-10
    s.free()
    fresh.free()


class Hierarchy:
    """amg_data built by the library's host set-up (the reference's own
    host structures), for the library-level entry points."""

    def __init__(self, net, smoother, omega, min_size=300):
        self.n = net.n
        self.keep = (net.ia, net.ja, net.val)
        self.A = L.CSR(net.n, net.n, net.nnz, ip(net.ia), ip(net.ja), dp(net.val))
        self.amg = L.AMGData()
        self.amg.threshold = 0.25
        self.amg.min_size = min_size
        self.amg.max_levels = 20
        self.amg.nu = 1
        self.amg.smoother = smoother
        self.amg.omega_jacobi = omega
        self.amg.coarsening = L.COARSEN_PMIS
        lib.amg_build_hierarchy(C.addressof(self.amg), C.byref(self.A))
        self.L = L.IcholFactor()
        lib.ichol_compute(C.byref(self.A), C.addressof(self.L))

    def free(self):
        lib.amg_free(C.addressof(self.amg))
        lib.ichol_free(C.addressof(self.L))


@pytest.mark.parametrize("name,code,omega", SMOOTHERS, ids=IDS)
def test_library_entry_points_match_the_oracle(net, oracle, name, code, omega):
    # amg_backslash / amg_fwdslash / amg_vcycle_precond / amg_ichol_precond
    # on the caller's amg_data, options as the reference's: IC-PCG coarsest
    # solve, exact sweeps
    case(f"library entry points on amg_data, smoother {name} omega {omega}")
    o = L.Options()
    lib.bn_options_default(C.byref(o))
    o.coarse_solver = L.COARSE_ICPCG
    o.gs_exact = 1
    lib.bn_set_library_options(C.byref(o))
    try:
        hier = Hierarchy(net, code, omega)
        h = oracle.build(net.ia, net.ja, net.val, coarsening="pmis", min_size=300, smoother=code, symmetrize=False)
        h.set_smoother(code, omega)
        rng = np.random.default_rng(7)
        g = rng.random(net.n) - 0.5
        x0 = rng.random(net.n)
        for nm, fn, hfn in (("backslash", lib.amg_backslash, oracle.lib.h_backslash),
                            ("fwdslash", lib.amg_fwdslash, oracle.lib.h_fwdslash)):
            x = x0.copy()
            fn(C.addressof(hier.amg), 0, dp(g), dp(x), 1)
            xr = x0.copy()
            hfn(h.h, 0, dp(g), dp(xr))
            e = rel(x, xr)
            print(f"amg_{nm} from a nonzero guess: rel err vs oracle {e:.2e}")
            assert e < 1e-10
        z = np.zeros(net.n)
        lib.amg_vcycle_precond(C.addressof(hier.amg), dp(g), dp(z))
        assert rel(z, h.vcycle(g)) < 1e-10
        lib.amg_ichol_precond(C.addressof(hier.amg), C.byref(hier.A), C.addressof(hier.L), dp(g), dp(z))
        assert rel(z, h.precond(g)) < 1e-10
        hier.free()
    finally:
        lib.bn_set_library_options(None)


# --- the MEX smoother's C-typed twin -------------------------------------
def oracle_sweep(oracle, n, ia, ja, val, b, x0, order1, nu, direction):
    f = oracle.lib.oracle_cf_gauss_seidel
    f.argtypes = [C.c_int, L.IP, L.IP, L.DP, L.DP, L.DP, L.DP, C.c_int, C.c_int, L.DP]
    x = np.zeros(n)
    f(n, ip(ia), ip(ja), dp(val), dp(b), dp(x0), dp(order1), nu, direction, dp(x))
    return x


def twin(n, ia, ja, val, b, x0, order1, nu, direction):
    x = np.zeros(n)
    rc = lib.cf_gauss_seidel(n, ip(ia), ip(ja), dp(val), dp(b), dp(x0), dp(order1), nu, direction, dp(x))
    assert rc == 0, L.last_error()
    return x


@pytest.mark.parametrize("nu", [1, 2])
@pytest.mark.parametrize("direction", [0, 1])
def test_cf_gauss_seidel_twin_matches_the_oracle(net, oracle, nu, direction):
    case(f"cf_gauss_seidel twin, nu {nu}, direction {direction}")
    s = AMGSolver(net.ia, net.ja, net.val, 0.25, 300, 20, 1, coarsening="pmis")
    lv = s.level(0)
    ia, ja, va = lv["A"]
    order1 = lv["cf_order"].astype(np.float64) + 1.0        # MATLAB's 1-based doubles
    n = len(ia) - 1
    rng = np.random.default_rng(11)
    b = rng.random(n)
    x0 = rng.random(n)
    x = twin(n, ia, ja, va, b, x0, order1, nu, direction)
    xo = oracle_sweep(oracle, n, ia, ja, va, b, x0, order1, nu, direction)
    e = rel(x, xo)
    print(f"rel err vs oracle {e:.2e}")
    assert e < 1e-13
    # a second level (denser Galerkin rows) and a random order
    lv1 = s.level(1)
    ia1, ja1, va1 = lv1["A"]
    n1 = len(ia1) - 1
    o1 = rng.permutation(n1).astype(np.float64) + 1.0
    b1 = rng.random(n1)
    x1 = twin(n1, ia1, ja1, va1, b1, np.zeros(n1), o1, nu, direction)
    assert rel(x1, oracle_sweep(oracle, n1, ia1, ja1, va1, b1, np.zeros(n1), o1, nu, direction)) < 1e-13
    s.free()


def test_cf_gauss_seidel_twin_closed_form():
    # one forward sweep in order pi: with B = A[pi, pi] = D + L + U (the
    # rows and columns in visiting order), the sweep solves
    # (D + L) x'[pi] = b[pi] - U x[pi]
    case("cf_gauss_seidel twin: dense closed form")
    rng = np.random.default_rng(12)
    n = 60
    M = sp.random(n, n, density=0.08, random_state=3)
    A = (M + M.T).tocsr()
    A.setdiag(np.abs(A).sum(axis=1).A.ravel() + 1.0)
    A.sort_indices()
    ia, ja, va = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    order = rng.permutation(n)
    b, x0 = rng.random(n), rng.random(n)
    x = twin(n, ia, ja, va, b, x0, order.astype(np.float64) + 1.0, 1, 0)
    Ap = A.toarray()[np.ix_(order, order)]
    lower = np.tril(Ap)
    upper = np.triu(Ap, 1)
    xp = np.linalg.solve(lower, b[order] - upper @ x0[order])
    ref = np.empty(n)
    ref[order] = xp
    assert rel(x, ref) < 1e-13
    # backward: the reverse order, i.e. the upper triangle of the permuted matrix
    xb = twin(n, ia, ja, va, b, x0, order.astype(np.float64) + 1.0, 1, 1)
    xq = np.linalg.solve(np.triu(Ap), b[order] - np.tril(Ap, -1) @ x0[order])
    refb = np.empty(n)
    refb[order] = xq
    assert rel(xb, refb) < 1e-13


def test_cf_gauss_seidel_device_form_equals_the_host_form(net):
    case("bn_cf_gauss_seidel_device vs cf_gauss_seidel")
    n = net.n
    rng = np.random.default_rng(13)
    order = rng.permutation(n).astype(np.int32)
    b, x0 = rng.random(n), rng.random(n)
    x = twin(n, net.ia, net.ja, net.val, b, x0, order.astype(np.float64) + 1.0, 2, 1)
    d = [L.DeviceArray(a) for a in (net.ia, net.ja, net.val, b, x0, order)]
    rc = lib.bn_cf_gauss_seidel_device(n, net.nnz, *[C.c_void_p(a.ptr) for a in d], 2, 1)
    assert rc == 0, L.last_error()
    xd = d[4].numpy()
    assert np.array_equal(x, xd)
    # a repeated row is rejected on the device path
    bad = order.copy()
    bad[0] = bad[1]
    d[5].upload(bad)
    assert lib.bn_cf_gauss_seidel_device(n, net.nnz, *[C.c_void_p(a.ptr) for a in d], 1, 0) == L.BN_ERR_ARG
    for a in d:
        a.free()
