"""Return conventions and argument checks at the C ABI that hold without a
GPU: the reference entry points keep the reference's returns
(C_amg1/src/amg_wrapper.c:148,198,214-225), the MEX twin rejects what the
MEX code would have read out of bounds, and the set-up no longer touches
the caller's rand() sequence."""
import ctypes as C

import numpy as np
import pytest

from blood_network_amd import _lib as L
from blood_network_amd import device_available, make_network

lib = L.lib
lib.rand.restype = C.c_int
lib.srand.argtypes = [C.c_uint]


@pytest.fixture(scope="module")
def small():
    return make_network(8, 5, seed=2)


def test_update_and_set_matrix_return_zero(small):
    net = small
    h = lib.amg_solver_create(net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(net.val), 0.25, 100, 20, 1)
    assert h
    v2 = net.val * 1.5
    assert lib.amg_solver_update(h, net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(v2)) == 0
    assert lib.amg_solver_set_matrix(h, net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(net.val)) == 0
    assert lib.amg_solver_size(h) == net.n
    lib.amg_solver_free(h)
    # a NULL handle is reported, not returned as a code the reference never returns
    assert lib.amg_solver_update(None, net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(v2)) == 0
    assert "not a solver handle" in L.last_error()


@pytest.mark.skipif(device_available(), reason="checks the no-device return")
def test_solve_without_a_device_returns_minus_maxiter(small):
    net = small
    h = lib.amg_solver_create(net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(net.val), 0.25, 100, 20, 1)
    x = np.full(net.n, 7.0)
    it = lib.amg_solver_solve(h, L.dp(net.b), L.dp(x), 1e-8, 123, 0)
    assert it == -123
    assert np.all(x == 7.0)              # x is left as it was
    assert "GPU" in L.last_error()
    lib.amg_solver_free(h)


def twin(n, ia, ja, val, b, x0, order, nu=1, direction=0):
    x = np.zeros(n)
    rc = lib.cf_gauss_seidel(n, L.ip(ia), L.ip(ja), L.dp(val), L.dp(b), L.dp(x0), L.dp(order), nu, direction,
                             L.dp(x))
    return rc, x


def test_cf_gauss_seidel_rejects_bad_arguments(small):
    net = small
    n = net.n
    b = np.ones(n)
    x0 = np.zeros(n)
    good = np.arange(1, n + 1, dtype=np.float64)
    for bad in (np.where(good == 5, 0.0, good),          # 0 is not a 1-based row number
                np.where(good == 5, n + 1.0, good),      # past the last row
                np.where(good == 5, 4.5, good)):         # not a row number at all
        rc, _ = twin(n, net.ia, net.ja, net.val, b, x0, bad)
        assert rc == L.BN_ERR_ARG and L.last_error().startswith("cf_gs:order")
    rc, _ = twin(n, net.ia, net.ja, net.val, b, x0, good, direction=2)
    assert rc == L.BN_ERR_ARG and L.last_error().startswith("cf_gs:args")
    rc, _ = twin(n, net.ia, net.ja, net.val, b, x0, good, nu=-1)
    assert rc == L.BN_ERR_ARG
    # a repeated row passes the range check and fails as a non-permutation
    # (on the device path, where the schedule is built)
    rep = good.copy()
    rep[3] = rep[4]
    rc, _ = twin(n, net.ia, net.ja, net.val, b, x0, rep)
    assert rc in (L.BN_ERR_ARG, L.BN_ERR_NO_DEVICE)
    # nu = 0 is x = x0 on any machine
    x0r = np.random.default_rng(0).random(n)
    rc, x = twin(n, net.ia, net.ja, net.val, b, x0r, good, nu=0)
    assert rc == 0 and np.array_equal(x, x0r)


def test_setup_leaves_the_callers_rand_sequence_alone(small):
    # PMIS draws its tie-break weights from a private random_r state with the
    # sequence srand(12345) gives; the caller's rand() stream is untouched
    net = small
    lib.srand(777)
    expect = [lib.rand() for _ in range(5)]
    lib.srand(777)
    first = [lib.rand() for _ in range(2)]
    h = lib.amg_solver_create(net.n, net.nnz, L.ip(net.ia), L.ip(net.ja), L.dp(net.val), 0.25, 100, 20, 1)
    rest = [lib.rand() for _ in range(3)]
    lib.amg_solver_free(h)
    assert first + rest == expect
