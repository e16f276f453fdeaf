"""The multicolour Gauss-Seidel sweep in one launch (bn_gs_walk) against
the same sweep as one launch per colour (k_rowsolve with the same lanes per
row, the C-then-F colour order of C_amg1/src/amg_cycle.c:52-80's smoother
as the multicolour recast runs it).  A row reads an earlier colour's
neighbours as the walk computes them and a later colour's from a copy of
the sweep's input, which is what the colour launches read, and forms the
same sums, so every PCG iterate and the solution must agree to the last
bit: FP64 and FP32 V-cycles, 1, 2, 4 and 8 lanes per row, one and two
sweeps per level, graph replay, and a walk that gives up and hands the
solve back to the colour launches.  The exact-order sweeps (natural order,
and the reference's cf_order) run the same way over their dependency
levels and must equal one launch per level to the last bit."""
import sys

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def net():
    return make_network(52, 10, seed=2)


def solver(monkeypatch, net, lanes, walk, nu=1, **kw):
    # the operator kernels untuned (their lanes, and so their sums, would
    # otherwise follow timings), the colour and one-launch lanes pinned
    monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
    monkeypatch.setenv("BN_GS_LANES", str(lanes))
    monkeypatch.setenv("BN_GS_WALK", str(lanes) if walk else "0")
    return AMGSolver(net.ia, net.ja, net.val, 0.25, 500, 20, nu, coarsening="pmis", **kw)


MODES = {"fp64": dict(), "fp32_fast": dict(precision="fp32", fine_ic="jacobi", ic_sweeps=4, max_colours=16,
                                            stop="relres")}


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("mode", list(MODES))
def test_walk_sweep_equals_colour_launches(monkeypatch, capfd, net, mode, lanes):
    case(f"GS sweep in one launch vs a launch per colour, {mode}, lanes {lanes}, n={net.n}")
    out = []
    for walk in (True, False):
        s = solver(monkeypatch, net, lanes, walk, **MODES[mode])
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)
        x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)   # replayed step graphs
        out.append((x, it, x2, it2))
        s.free()
    (xw, iw, xw2, iw2), (xl, il, xl2, il2) = out
    print(f"iterations walk {iw}/{iw2}, colour launches {il}/{il2}; entries differing {np.sum(xw != xl)}, "
          f"max rel {np.max(np.abs(xw - xl)) / np.max(np.abs(xl)):.2e}; walk repeat differs {np.sum(xw != xw2)}",
          flush=True)
    assert "timed out" not in capfd.readouterr().err  # the walk itself ran, no hand-back
    assert iw > 0 and iw == il == iw2 == il2
    assert np.array_equal(xw, xl) and np.array_equal(xw2, xl2) and np.array_equal(xw, xw2)


def test_walk_two_sweeps_per_level(monkeypatch, capfd, net):
    case(f"GS sweep in one launch, nu = 2, n={net.n}")
    res = []
    for walk in (True, False):
        s = solver(monkeypatch, net, 4, walk, nu=2)
        res.append(s.solve(net.b, tol=1e-10, maxiter=500))
        s.free()
    (xw, iw), (xl, il) = res
    print(f"iterations {iw} / {il}", flush=True)
    assert "timed out" not in capfd.readouterr().err
    assert iw == il and np.array_equal(xw, xl)


def test_walk_sweep_that_gives_up_hands_back(monkeypatch, capfd, net):
    case(f"GS sweep in one launch forced to give up (BN_TRI_WALK_LIMIT=-1), n={net.n}")
    ref = solver(monkeypatch, net, 2, False, **MODES["fp32_fast"])
    x0, i0 = ref.solve(net.b, tol=1e-10, maxiter=500)
    ref.free()
    monkeypatch.setenv("BN_TRI_WALK_LIMIT", "-1")
    s = solver(monkeypatch, net, 2, True, **MODES["fp32_fast"])
    x1, i1 = s.solve(net.b, tol=1e-10, maxiter=500)
    x2, i2 = s.solve(net.b, tol=1e-10, maxiter=500)
    s.free()
    err = capfd.readouterr().err
    print(f"iterations {i1}/{i2} vs {i0}", flush=True)
    assert "timed out" in err
    assert i1 == i2 == i0 and np.array_equal(x1, x0) and np.array_equal(x2, x0)


# ---------------------------------------------------------------------------
# exact-order sweeps (SMOOTHER_GS natural order, SMOOTHER_CF_GS with gs_exact:
# C_amg1/src/amg_cycle.c:21-80) in one launch per sweep (bn_sweep_parts)
# against one launch per dependency level

EXACT = {"gs_natural": dict(smoother="gs", coarse="icpcg"), "cf_exact": dict(gs_exact=True, coarse="icpcg")}


def exact_solver(monkeypatch, net, walk, parts=None, **kw):
    monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
    monkeypatch.setenv("BN_SWEEP_WALK", "1" if walk else "0")
    if parts:
        monkeypatch.setenv("BN_SWEEP_PARTS", str(parts))
    else:
        monkeypatch.delenv("BN_SWEEP_PARTS", raising=False)
    return AMGSolver(net.ia, net.ja, net.val, 0.25, 500, 20, 1, coarsening="pmis", **kw)


@pytest.fixture(scope="module")
def net_small():
    return make_network(30, 9, seed=4)


@pytest.mark.parametrize("parts", [None, 1, 64])
@pytest.mark.parametrize("kind", list(EXACT))
def test_exact_sweep_in_one_launch_equals_level_launches(monkeypatch, capfd, net_small, kind, parts):
    case(f"exact {kind} sweeps in one launch (parts {parts or 'default'}) vs a launch per level, n={net_small.n}")
    res = []
    for walk in (True, False):
        s = exact_solver(monkeypatch, net_small, walk, parts, **EXACT[kind])
        x, it = s.solve(net_small.b, tol=1e-10, maxiter=500)
        res.append((x, it))
        s.free()
    (xw, iw), (xl, il) = res
    print(f"iterations {iw} / {il}; entries differing {np.sum(xw != xl)}", flush=True)
    assert "timed out" not in capfd.readouterr().err  # the walk itself ran, no hand-back
    assert iw > 0 and iw == il and np.array_equal(xw, xl)


def test_exact_sweep_that_gives_up_hands_back(monkeypatch, capfd, net_small):
    case(f"exact cf sweep in one launch forced to give up (BN_TRI_WALK_LIMIT=-1), n={net_small.n}")
    ref = exact_solver(monkeypatch, net_small, False, **EXACT["cf_exact"])
    x0, i0 = ref.solve(net_small.b, tol=1e-10, maxiter=500)
    ref.free()
    monkeypatch.setenv("BN_TRI_WALK_LIMIT", "-1")
    s = exact_solver(monkeypatch, net_small, True, 4, **EXACT["cf_exact"])
    x1, i1 = s.solve(net_small.b, tol=1e-10, maxiter=500)
    s.free()
    err = capfd.readouterr().err
    print(f"iterations {i1} vs {i0}", flush=True)
    assert "timed out" in err
    assert i1 == i0 and np.array_equal(x1, x0)
