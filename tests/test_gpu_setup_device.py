"""The per-row set-up stages on the device (setup_dev.hip) against the host
routines and the oracle: strength (C_amg1/src/rs_classical.c:30-88, and
the MEX rule), the PMIS rounds (rs_classical.c:246-371) and standard
interpolation (rs_classical.c:643-805).  Every level's S, split, cf_order,
P, operator and both IC(0) factors must equal, bit for bit, the hierarchy
built with BN_DEVICE_SETUP=off and the oracle's; BN_SETUP_LOG shows which
levels ran the device stages, so each case checks that they did."""
import sys

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network, workload_network
from test_oracle_pins import random_graph_laplacian

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def same(a, b, tag=""):
    assert a.num_levels() == b.num_levels()
    for lev in range(a.num_levels()):
        la, lb = a.level(lev), b.level(lev)
        assert la.keys() == lb.keys()
        for k in lb:
            if isinstance(lb[k], tuple):
                for x, y in zip(la[k], lb[k]):
                    assert np.array_equal(x, y), (tag, lev, k)
            else:
                assert np.array_equal(la[k], lb[k]), (tag, lev, k)
    for x, y in zip(a.ichol(0), b.ichol(0)):
        assert np.array_equal(x, y), tag


def build(ia, ja, val, device, monkeypatch, capfd, min_nnz="0", **kw):
    """device: the per-row stages on the device for levels of at least
    min_nnz operator entries, whatever the coarsening; False: all on the
    host; None: the automatic choice.  Returns the solver and the number of
    levels that ran the device stages."""
    monkeypatch.setenv("BN_SETUP_LOG", "1")
    if device is None:
        monkeypatch.delenv("BN_DEVICE_SETUP", raising=False)
    elif device:
        monkeypatch.setenv("BN_DEVICE_SETUP", "all")
        monkeypatch.setenv("BN_DEVICE_SETUP_MIN", min_nnz)
    else:
        monkeypatch.setenv("BN_DEVICE_SETUP", "off")
    capfd.readouterr()
    s = AMGSolver(ia, ja, val, **kw)
    err = capfd.readouterr().err
    return s, err.count("device stages")


@pytest.mark.parametrize("coarsening", ["rs", "pmis", "hmis"])
@pytest.mark.parametrize("strength", ["c", "mex"])
def test_device_stages_equal_host_small(coarsening, strength, monkeypatch, capfd):
    case(f"device set-up stages vs host, 24^3 lattice + trees, {coarsening}, strength rule {strength}")
    net = make_network(24, 10, seed=5)
    kw = dict(coarsening=coarsening, min_size=100, strength=strength)
    d, ndev = build(net.ia, net.ja, net.val, True, monkeypatch, capfd, **kw)
    h, nh = build(net.ia, net.ja, net.val, False, monkeypatch, capfd, **kw)
    print(f"levels {d.num_levels()}, device-stage levels {ndev}", flush=True)
    assert d.num_levels() >= 4 and ndev == d.num_levels() - 1 and nh == 0
    same(d, h, coarsening)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_stages_equal_oracle_random_graphs(seed, monkeypatch, capfd):
    case(f"device set-up stages vs the oracle, random graph Laplacian seed {seed}")
    import oracle_lib
    A = random_graph_laplacian(900, 7, seed=seed)
    for c in ("rs", "pmis", "hmis"):
        d, ndev = build(A.indptr, A.indices, A.data, True, monkeypatch, capfd, coarsening=c, min_size=50)
        o = oracle_lib.oracle().build(A.indptr, A.indices, A.data, coarsening=c, min_size=50)
        assert ndev >= 1
        assert d.num_levels() == o.num_levels()
        for lev in range(d.num_levels()):
            for k, v in o.level(lev).items():
                w = d.level(lev)[k]
                if isinstance(v, tuple):
                    for x, y in zip(w, v):
                        assert np.array_equal(x, y), (c, lev, k)
                else:
                    assert np.array_equal(w, v), (c, lev, k)


def test_device_stages_default_threshold_1m_network(monkeypatch, capfd):
    """The 1.07M-unknown network with the automatic choice (PMIS: its two
    levels of more than 4M entries run the device stages) against the host
    set-up, and the same solve from both hierarchies."""
    net = workload_network("lattice100_trees15")
    case(f"device set-up stages, automatic choice, lattice100_trees15 n={net.n}, PMIS")
    # kernel choices pinned (the upload at creation skips the timing-based
    # tuner), so the two solvers run the same kernels
    monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
    d, ndev = build(net.ia, net.ja, net.val, None, monkeypatch, capfd, coarsening="pmis")
    h, _ = build(net.ia, net.ja, net.val, False, monkeypatch, capfd, coarsening="pmis")
    print(f"levels {d.num_levels()}, device-stage levels {ndev}", flush=True)
    assert ndev == 2
    same(d, h, "lattice100_trees15")
    x1, it1 = d.solve(net.b, tol=1e-10)
    x2, it2 = h.solve(net.b, tol=1e-10)
    assert it1 == it2 and np.array_equal(x1, x2)
