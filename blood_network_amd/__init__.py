"""MI355X-native AMG-preconditioned PCG for vessel-network pressure systems.

The numerics live in ``libbnamg.so`` (host C set-up + gfx950 HIP kernels,
C ABI in ``include/amg.h`` and ``include/bn_amg_ext.h``).  Importing the
package loads that library and fails if it is missing.
"""
from ._lib import device_available, last_error, lib
from .amg import AMGSolver, SolveReport
from .inverse import PressureFit
from .netgen import WORKLOADS, VesselNetwork, make_network, uneven_batch, union_networks

__all__ = ["AMGSolver", "SolveReport", "PressureFit", "VesselNetwork", "make_network", "WORKLOADS", "union_networks", "uneven_batch",
           "device_available", "last_error", "lib"]
