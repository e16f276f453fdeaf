"""MI355X-native AMG-preconditioned PCG for vessel-network pressure systems.

The numerics live in ``libbnamg.so`` (host C set-up + gfx950 HIP kernels,
C ABI in ``include/amg.h`` and ``include/bn_amg_ext.h``).  Importing the
package loads that library and fails if it is missing.
"""
from ._lib import device_available, last_error, lib
from .amg import AMGSolver, SolveReport
from .inverse import PressureFit
from .netgen import (WORKLOAD_OPTIONS, WORKLOADS, NetworkCase, VesselNetwork, make_network, read_network_case,
                     uneven_batch, union_networks, workload_network, write_labelled, write_pressures)

__all__ = ["AMGSolver", "SolveReport", "PressureFit", "VesselNetwork", "make_network", "workload_network", "WORKLOADS",
           "WORKLOAD_OPTIONS", "union_networks", "uneven_batch",
           "NetworkCase", "read_network_case", "write_pressures", "write_labelled",
           "device_available", "last_error", "lib"]
