"""The torch.autograd.Function wrapper of the pressure solve and its
adjoint (tests/autograd_harness.py), run in a process of its own that
imports torch before the package so both share one HIP runtime."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.timeout(900)
def test_autograd_function_matches_adjoint_and_direct_gradient():
    pytest.importorskip("torch")
    print("\n[case] torch.autograd.Function harness (own process)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "autograd_harness.py")], cwd=os.path.dirname(HERE),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=850)
    print(r.stdout[-4000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    res = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])
    assert res["ok"]
    assert res["cpu_vs_direct"] < 1e-6
    assert res["cuda_vs_cpu"] < 1e-9
