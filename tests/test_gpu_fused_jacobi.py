"""Jacobi-smoothed levels with their steps fused into the launches around
them against the same V-cycle with separate launches (BN_FUSE_JACOBI=0).
The first step from zero (C_amg1/src/amg_cycle.c:85-129's x = w b with a
zero start) is written by the restriction that forms the level's b, and
each further step (the residual and the update) is one launch that writes
the other of two buffers; the values are those of the separate launches,
so every PCG iterate and the solution must agree to the last bit: the fast
FP32 path (Jacobi below the Gauss-Seidel levels), l1-Jacobi and damped
Jacobi on every level in FP64 and FP32, one to three sweeps per level, and
graph replay."""
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


MODES = {
    "fp32_fast": dict(coarsening="hmis", precision="fp32", fine_ic="jacobi", ic_sweeps=4, max_colours=16,
                      stop="relres"),
    "fp64_l1": dict(coarsening="pmis", smoother="l1"),
    "fp32_l1": dict(coarsening="pmis", smoother="l1", precision="fp32", fine_ic="jacobi", ic_sweeps=3,
                    stop="relres"),
    "fp64_jacobi": dict(coarsening="pmis", smoother="jacobi"),
}


@pytest.mark.parametrize("nu", [1, 2, 3])
@pytest.mark.parametrize("mode", list(MODES))
def test_fused_jacobi_equals_separate_launches(monkeypatch, net, mode, nu):
    case(f"fused Jacobi steps vs separate launches, {mode}, nu {nu}, n={net.n}")
    out = []
    for fuse in ("1", "0"):
        # untuned operator kernels, so both solvers run the same ones
        monkeypatch.setenv("BN_NO_AUTOTUNE", "1")
        monkeypatch.setenv("BN_FUSE_JACOBI", fuse)
        s = AMGSolver(net.ia, net.ja, net.val, 0.25, 500, 20, nu, coarse="dense", **MODES[mode])
        if mode == "fp64_jacobi":
            s.set_smoother("jacobi", 0.7)
        x, it = s.solve(net.b, tol=1e-10, maxiter=500)
        x2, it2 = s.solve(net.b, tol=1e-10, maxiter=500)   # replayed step graphs
        out.append((x, it, x2, it2, s.num_levels()))
        s.free()
    (xf, itf, xf2, itf2, nl), (xs, its, xs2, its2, _) = out
    print(f"levels {nl}; iterations fused {itf}/{itf2}, separate {its}/{its2}; entries differing "
          f"{np.sum(xf != xs)}", flush=True)
    assert nl >= 4
    assert itf > 0 and itf == its == itf2 == its2
    assert np.array_equal(xf, xs) and np.array_equal(xf2, xs2) and np.array_equal(xf, xf2)
