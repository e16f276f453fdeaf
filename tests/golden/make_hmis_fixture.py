"""Writes tests/golden/hmis_hub_level.npz: the strength matrix S of the
7th level (16,782 rows) of the 160^3-lattice + depth-17-trees network under
HMIS coarsening (theta 0.25).  On this S the C_amg1 HMIS greedy pass lifts
a node's lambda above the bucket array it allocates (max_lam + longest S
row + 1, C_amg1/src/rs_classical.c:410-418), so the compiled reference
writes past that array; the oracle and the library size their bucket
arrays for every reachable lambda.  The fixture pins the split both give.

Run from the repository root after building the library:
    python tests/golden/make_hmis_fixture.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from blood_network_amd import AMGSolver, make_network  # noqa: E402

net = make_network(160, 17)
s = AMGSolver(net.ia, net.ja, net.val, coarsening="hmis")
lv = s.level(5)
ia, ja, val = lv["S"]
cf = lv["mis"]
np.savez_compressed(os.path.join(ROOT, "tests", "golden", "hmis_hub_level.npz"),
                    s_ia=ia, s_ja=ja, s_val=val, mis=cf)
print("rows", len(ia) - 1, "nnz(S)", len(ja), "C points", int(cf.sum()))
