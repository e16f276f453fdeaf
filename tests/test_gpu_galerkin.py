"""The set-up's Galerkin product on the device (galerkin.hip) against the
host product (setup.c's dcsr_rap and (A + A^T)/2, C_amg1/src/sparse_ops.c
and build_hierarchy.c semantics): every level's operator, interpolation,
split and IC(0) factor bit for bit, against the hierarchy built with
BN_HOST_GALERKIN=1 and against the oracle's host restatement."""
import sys
import time

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network
from test_oracle_pins import random_graph_laplacian

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def same(a, b):
    assert a.num_levels() == b.num_levels()
    for lev in range(a.num_levels()):
        la, lb = a.level(lev), b.level(lev)
        assert la.keys() == lb.keys()
        for k in lb:
            if isinstance(lb[k], tuple):
                for x, y in zip(la[k], lb[k]):
                    assert np.array_equal(x, y), (lev, k)
            else:
                assert np.array_equal(la[k], lb[k]), (lev, k)
    for x, y in zip(a.ichol(0), b.ichol(0)):
        assert np.array_equal(x, y)


def build(ia, ja, val, host, monkeypatch, **kw):
    """host: the product on the host; otherwise on the device at every
    level (BN_DEVICE_GALERKIN_MIN=0 lifts the size floor)"""
    monkeypatch.setenv("BN_DEVICE_GALERKIN_MIN", "0")
    if host:
        monkeypatch.setenv("BN_HOST_GALERKIN", "1")
    else:
        monkeypatch.delenv("BN_HOST_GALERKIN", raising=False)
    t0 = time.perf_counter()
    s = AMGSolver(ia, ja, val, **kw)
    return s, time.perf_counter() - t0


@pytest.mark.parametrize("coarsening", ["rs", "pmis", "hmis"])
def test_device_product_equals_host_small(coarsening, monkeypatch):
    case(f"device Galerkin product vs host, 24^3 lattice + trees, {coarsening}")
    net = make_network(24, 10, seed=5)
    d, _ = build(net.ia, net.ja, net.val, False, monkeypatch, coarsening=coarsening, min_size=100)
    h, _ = build(net.ia, net.ja, net.val, True, monkeypatch, coarsening=coarsening, min_size=100)
    assert d.num_levels() >= 4
    same(d, h)
    x1, it1 = d.solve(net.b, tol=1e-10)
    x2, it2 = h.solve(net.b, tol=1e-10)
    print(f"levels {d.num_levels()}, iterations {it1} / {it2}")
    assert it1 == it2


@pytest.mark.parametrize("seed", [0, 1])
def test_device_product_equals_oracle_random_graphs(seed, monkeypatch):
    case(f"device Galerkin product vs the oracle, random graph Laplacian seed {seed}")
    import oracle_lib
    A = random_graph_laplacian(900, 7, seed=seed)
    for c in ("rs", "pmis", "hmis"):
        d, _ = build(A.indptr, A.indices, A.data, False, monkeypatch, coarsening=c, min_size=50)
        o = oracle_lib.oracle().build(A.indptr, A.indices, A.data, coarsening=c, min_size=50)
        assert d.num_levels() == o.num_levels()
        for lev in range(d.num_levels()):
            for k, v in o.level(lev).items():
                w = d.level(lev)[k]
                if isinstance(v, tuple):
                    for x, y in zip(w, v):
                        assert np.array_equal(x, y), (c, lev, k)
                else:
                    assert np.array_equal(w, v), (c, lev, k)
        d.free()


@pytest.mark.parametrize("coarsening", ["pmis", "hmis"])
def test_device_product_equals_host_full_network(coarsening, monkeypatch):
    """1,065,532 unknowns: hub rows whose Galerkin entries take ~10^5 terms
    (one wavefront per long entry)."""
    case(f"device Galerkin product vs host, 100^3 lattice + depth-15 trees, {coarsening}")
    monkeypatch.setenv("BN_SETUP_LOG", "1")
    net = make_network(100, 15, 1, seed=0)
    d, td = build(net.ia, net.ja, net.val, False, monkeypatch, coarsening=coarsening)
    h, th = build(net.ia, net.ja, net.val, True, monkeypatch, coarsening=coarsening)
    print(f"set-up: device product {td:.2f} s, host product {th:.2f} s; levels {d.num_levels()}", flush=True)
    same(d, h)
    d.free()
    h.free()
