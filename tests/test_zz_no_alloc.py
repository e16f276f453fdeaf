"""bn_solver_adjoint_grad_noalloc, the allocation-free form of
bn_solver_adjoint_grad (static device scratch, no hipMalloc), against that
entry point through the C ABI: bit for bit, on two calls in a row (the
second smaller), with the segment list long enough to pass through the
scratch in several chunks, and its size limit."""
import sys

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network
from blood_network_amd import _lib as L

pytestmark = pytest.mark.gpu
SCRATCH_EDGES = 1 << 20
SCRATCH_DIRICHLET = 1 << 16


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    net = make_network(16, 9, seed=2)
    s = AMGSolver(net.ia, net.ja, net.val, min_size=500)
    p, _ = s.solve(net.b, tol=1e-12, maxiter=500)
    yield net, s, p
    s.free()


def both(s, p, dLdp, eu, ev, g, pdir):
    a = s.adjoint_grad(p, dLdp, eu, ev, g, pdir, tol=1e-11)
    b = s.adjoint_grad_noalloc(p, dLdp, eu, ev, g, pdir, tol=1e-11)
    return a, b


def same(a, b):
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)
    assert a[3] == b[3] and a[3] > 0


def test_noalloc_equals_the_allocating_entry_on_two_calls(setup):
    net, s, p = setup
    rng = np.random.default_rng(1)
    case(f"call 1: all {net.n_edges} segments")
    dLdp = rng.normal(size=net.n)
    same(*both(s, p, dLdp, net.eu, net.ev, net.g, net.p_dirichlet))
    case("call 2: a smaller segment list and another dL/dp")
    m2 = net.n_edges // 3
    sel = rng.choice(net.n_edges, size=m2, replace=False)
    dLdp2 = np.zeros(net.n)
    dLdp2[rng.choice(net.n, 50, replace=False)] = 1.0
    same(*both(s, p, dLdp2, net.eu[sel], net.ev[sel], net.g[sel], net.p_dirichlet))


def test_noalloc_passes_long_segment_lists_in_chunks(setup):
    net, s, p = setup
    reps = SCRATCH_EDGES * 2 // net.n_edges + 1
    eu, ev, g = np.tile(net.eu, reps), np.tile(net.ev, reps), np.tile(net.g, reps)
    case(f"{len(eu)} segments: {-(-len(eu) // SCRATCH_EDGES)} passes through the scratch")
    assert len(eu) > 2 * SCRATCH_EDGES
    dLdp = np.random.default_rng(2).normal(size=net.n)
    a, b = both(s, p, dLdp, eu, ev, g, net.p_dirichlet)
    same(a, b)
    # the tiled list repeats the first copy's gradient exactly
    m = net.n_edges
    assert np.array_equal(b[2][:m], b[2][(reps - 1) * m:reps * m])


def test_noalloc_rejects_more_dirichlet_pressures_than_it_holds(setup):
    net, s, p = setup
    case("more Dirichlet pressures than the scratch holds")
    eu = net.eu.copy()
    eu[0] = -(SCRATCH_DIRICHLET + 1)
    pdir = np.zeros(SCRATCH_DIRICHLET + 1)
    with pytest.raises(RuntimeError, match="Dirichlet"):
        s.adjoint_grad_noalloc(p, np.ones(net.n), eu, net.ev, net.g, pdir)
    # and the solver still works afterwards
    same(*both(s, p, np.ones(net.n), net.eu, net.ev, net.g, net.p_dirichlet))
