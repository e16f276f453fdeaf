"""The kernel census (bn_solver_kernel_census): its solve is the ordinary
solve, and its per-symbol algorithmic bytes follow the per-launch formulas
(each array touched once) for the launches a V-cycle makes."""
import sys

import numpy as np
import pytest

from blood_network_amd import AMGSolver, make_network
from blood_network_amd import _lib as L

pytestmark = pytest.mark.gpu


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def test_census_solve_and_byte_counts():
    case("kernel census on a 20^3 network, FP32 V-cycle")
    net = make_network(20, 9, seed=1)
    kw = dict(coarsening="hmis", min_size=300, precision="fp32", fine_ic="jacobi", ic_sweeps=4, max_colours=16,
              stop="relres")
    s = AMGSolver(net.ia, net.ja, net.val, **kw)
    db, dx = L.DeviceArray(net.b), L.DeviceArray(np.zeros(net.n))
    it = s.solve_device(db.ptr, dx.ptr, tol=1e-8)
    x_ref = dx.numpy()
    dx.upload(np.zeros(net.n))
    stats = s.kernel_census(db.ptr, dx.ptr, tol=1e-8, reps=2)
    x = dx.numpy()
    # the census solve is the solve (eager launches, the same kernels)
    assert np.array_equal(x, x_ref)
    names = {c["name"]: c for c in stats}
    print("\n".join(f"{c['name']}: {c['launches']} launches, {c['bytes'] / 1e6:.2f} MB, "
                    f"{c['us_solve']:.1f} us as run, {c['us_cold']:.1f} us cold" for c in stats[:8]))
    assert all(c["launches"] > 0 and c["bytes"] > 0 and c["us_solve"] > 0 and c["us_cold"] > 0 for c in stats)
    # the outer FP64 product: one launch per PCG step plus the initial
    # residual, each moving nnz*12 + (n+1)*4 + n*8 (x) + n*8 (y) [+ n*8 (b)]
    spmv = [c for c in stats if c["name"].startswith(("k_csr_blocks<double", "k_csr<double"))]
    assert sum(c["launches"] for c in spmv) == it + 1
    nnz, n = net.nnz, net.n
    per = nnz * 12.0 + (n + 1) * 4.0 + 2 * n * 8.0
    assert abs(sum(c["bytes"] for c in spmv) - (it * per + per + n * 8.0)) < 1.0
    # the IC(0) Jacobi sweeps over the fine lower/upper factors: 4 + 4 per
    # application, less the upper triangle's first step (the last lower sweep
    # writes it) and, when the fine residual runs the row-block kernel, the
    # lower triangle's first step (that residual writes it)
    jac = sum(c["launches"] for c in stats if c["name"].startswith("k_jacobi_rows"))
    fused = names.get("k_csr_blocks<float, 1, true>", {}).get("launches", 0)
    assert fused in (0, it + 1)
    assert jac == (6 if fused else 7) * (it + 1)
    s.free()
