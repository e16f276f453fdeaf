"""Synthetic vessel networks (wrapper over bn_netgen_* in libbnamg.so).

``make_network(L, depth, n_trees=1, seed=0)`` returns a ``VesselNetwork``
with the SPD pressure system A p = b (Dirichlet roots eliminated) and the
per-segment data the adjoint needs.  The named workloads below are the
four sizes the benchmarks and parity tests use.
"""
import ctypes as C
import hashlib

import numpy as np

from . import _lib as L

__all__ = ["VesselNetwork", "make_network", "workload_network", "WORKLOADS", "WORKLOAD_OPTIONS",
           "union_networks", "uneven_batch", "NetworkCase",
           "read_network_case", "write_pressures", "write_labelled"]

# name -> (lattice edge, tree depth, trees per kind, coarsening)
WORKLOADS = {
    "lattice16_trees9": (16, 9, 1, "rs"),        # 5,116 unknowns
    "lattice100_trees15": (100, 15, 1, "pmis"),  # 1,065,532 unknowns
    "lattice160_trees17": (160, 17, 1, "hmis"),  # 4,358,140 unknowns
    "lattice100_16x16trees11": (100, 11, 16, "pmis"),  # 1,065,472 unknowns
}
# generator options beyond (lattice edge, depth, trees): the many-tree
# network roots each tree at a random boundary-face node, its leaves
# attached to the nearest free lattice nodes
WORKLOAD_OPTIONS = {
    "lattice100_16x16trees11": {"boundary_roots": 1},
}


class VesselNetwork:
    def __init__(self, net, params):
        self.params = params
        self.n = net.n
        self.nnz = net.nnz
        self.n_edges = net.n_edges
        self.n_nodes = net.n_nodes
        self.n_dirichlet = net.n_dirichlet
        arr = np.ctypeslib.as_array
        self.ia = arr(net.ia, (net.n + 1,)).copy()
        self.ja = arr(net.ja, (net.nnz,)).copy()
        self.val = arr(net.val, (net.nnz,)).copy()
        self.b = arr(net.b, (net.n,)).copy()
        m = net.n_edges
        self.eu = arr(net.eu, (m,)).copy()
        self.ev = arr(net.ev, (m,)).copy()
        self.radius = arr(net.radius, (m,)).copy()
        self.length = arr(net.length, (m,)).copy()
        self.g = arr(net.g, (m,)).copy()
        self.pos_uv = arr(net.pos_uv, (m,)).copy()
        self.pos_vu = arr(net.pos_vu, (m,)).copy()
        self.diag_pos = arr(net.diag_pos, (net.n,)).copy()
        self.p_dirichlet = arr(net.p_dirichlet, (net.n_dirichlet,)).copy()
        self.dirichlet_is_art = arr(net.dirichlet_is_art, (net.n_dirichlet,)).copy()
        self.mu = params.mu

    def conductance(self, radius):
        """g = pi r^4 / (8 mu len) per segment."""
        r = np.asarray(radius, dtype=np.float64)
        return np.pi / (8.0 * self.mu) * (r * r) * (r * r) / self.length

    def assemble(self, g, p_dir=None):
        """(val, b) of A and b for segment conductances g on the fixed pattern."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        val = np.zeros(self.nnz)
        b = np.zeros(self.n)
        net = self._struct()
        pd = None if p_dir is None else np.ascontiguousarray(p_dir, dtype=np.float64)
        L.lib.bn_netgen_assemble(C.byref(net), L.dp(g), L.dp(pd) if pd is not None else None,
                                 L.dp(val), L.dp(b))
        return val, b

    def rhs_for(self, p_dir):
        """b for other Dirichlet pressures (A does not change)."""
        return self.assemble(self.g, p_dir)[1]

    def sha256(self):
        h = hashlib.sha256()
        for a in (self.ia, self.ja, self.val, self.b):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    def _struct(self):
        # a bn_network view over this object's arrays (assemble only reads)
        s = L.Network()
        s.n_nodes, s.n, s.n_edges, s.nnz, s.n_dirichlet = (self.n_nodes, self.n, self.n_edges,
                                                           self.nnz, self.n_dirichlet)
        s.ia, s.ja, s.val, s.b = L.ip(self.ia), L.ip(self.ja), L.dp(self.val), L.dp(self.b)
        s.eu, s.ev = L.ip(self.eu), L.ip(self.ev)
        s.radius, s.length, s.g = L.dp(self.radius), L.dp(self.length), L.dp(self.g)
        s.pos_uv, s.pos_vu, s.diag_pos = L.ip(self.pos_uv), L.ip(self.pos_vu), L.ip(self.diag_pos)
        s.p_dirichlet = L.dp(self.p_dirichlet)
        s.dirichlet_is_art = L.ip(self.dirichlet_is_art)
        s.node_unknown = None
        self._keep = s
        return s


def make_network(L_edge, depth, n_trees=1, seed=0, **overrides):
    p = L.NetParams()
    L.lib.bn_netgen_default_params(C.byref(p))
    p.L, p.depth, p.n_trees, p.seed = L_edge, depth, n_trees, seed
    for k, v in overrides.items():
        setattr(p, k, v)
    net = L.Network()
    rc = L.lib.bn_netgen_build(C.byref(p), C.byref(net))
    if rc != 0:
        raise ValueError(f"bad network parameters L={L_edge} depth={depth} n_trees={n_trees}")
    try:
        return VesselNetwork(net, p)
    finally:
        L.lib.bn_netgen_free(C.byref(net))


def workload_network(name, seed=0):
    """The named workload's network (WORKLOADS, WORKLOAD_OPTIONS)."""
    Ledge, depth, ntrees, _ = WORKLOADS[name]
    return make_network(Ledge, depth, ntrees, seed=seed, **WORKLOAD_OPTIONS.get(name, {}))


def union_networks(nets, rhs=None):
    """Stack independent networks into one block-diagonal system.

    Returns (ia, ja, val, b, seg_ptr): the union CSR (rows of network g
    are seg_ptr[g] .. seg_ptr[g+1]-1, columns shifted by the same offset,
    no coupling between networks) and its right-hand side (each network's
    own b, or rhs[g] when given)."""
    sizes = np.array([net.n for net in nets], dtype=np.int64)
    seg_ptr = np.zeros(len(nets) + 1, dtype=np.int64)
    seg_ptr[1:] = np.cumsum(sizes)
    nnz = np.array([net.nnz for net in nets], dtype=np.int64)
    if seg_ptr[-1] >= 2**31 or nnz.sum() >= 2**31:
        raise ValueError("union too large for int32 CSR")
    ia = np.zeros(seg_ptr[-1] + 1, dtype=np.int32)
    ja = np.empty(nnz.sum(), dtype=np.int32)
    val = np.empty(nnz.sum())
    b = np.empty(seg_ptr[-1])
    q = 0
    for g, net in enumerate(nets):
        r0, r1 = seg_ptr[g], seg_ptr[g + 1]
        ia[r0 + 1:r1 + 1] = net.ia[1:] + q
        ja[q:q + net.nnz] = net.ja + r0
        val[q:q + net.nnz] = net.val
        b[r0:r1] = net.b if rhs is None else rhs[g]
        q += net.nnz
    return ia, ja, val, b, seg_ptr.astype(np.int32)


def uneven_batch(count=64, seed=0, edge=(12, 40), depth=(6, 10)):
    """`count` independent networks with lattice edge ~ U{edge} and tree
    depth ~ U{depth} (one arterial and one venous tree each), network g
    generated with seed + g."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(edge[0], edge[1] + 1, count)
    ds = rng.integers(depth[0], depth[1] + 1, count)
    nets = []
    for g, (Le, d) in enumerate(zip(Ls, ds)):
        # the leaves (2 * 2^(d-1)) must fit on distinct lattice nodes
        while 2 * 2 ** (int(d) - 1) > int(Le) ** 3:
            d -= 1
        nets.append(make_network(int(Le), int(d), 1, seed=seed + g))
    return nets


class NetworkCase:
    """A vessel-network case read from <base>.fMx/.pMx/.dia/.BC and its
    interior pressure system (libbnamg bn_nwk_read / bn_nwk_system, after
    Matlab_amg/caseReaderMJ.m and read_nwk.m).  Indices are 0-based except
    ``interior``, which keeps the files' 1-based node numbers."""

    def __init__(self, base):
        w = L.Nwk()
        if L.lib.bn_nwk_read(str(base).encode(), C.byref(w)):
            raise ValueError(L.last_error())
        try:
            self.n_points, self.n_faces, self.n_boundary = w.np, w.nf, w.nbc
            self.points = np.ctypeslib.as_array(w.pts, shape=(w.np, 3)).copy()
            self.faces = np.ctypeslib.as_array(w.face, shape=(w.nf, 2)).copy()
            self.diameters = (np.ctypeslib.as_array(w.dia, shape=(w.nf,)).copy() if bool(w.dia) else None)
            nb = max(w.nbc, 1)
            self.bc_nodes = np.ctypeslib.as_array(w.bc_node, shape=(nb,)).copy()[:w.nbc]
            self.bc_pressures = np.ctypeslib.as_array(w.bc_val, shape=(nb,)).copy()[:w.nbc]
            sysm = L.NwkSys()
            if L.lib.bn_nwk_system(C.byref(w), C.byref(sysm)):
                raise ValueError(L.last_error())
            try:
                n, nnz = sysm.n, sysm.nnz
                self.n = n
                self.ia = np.ctypeslib.as_array(sysm.ia, shape=(n + 1,)).copy()
                self.ja = np.ctypeslib.as_array(sysm.ja, shape=(max(nnz, 1),)).copy()[:nnz]
                self.val = np.ctypeslib.as_array(sysm.val, shape=(max(nnz, 1),)).copy()[:nnz]
                self.b = np.ctypeslib.as_array(sysm.b, shape=(max(n, 1),)).copy()[:n]
                self.interior = np.ctypeslib.as_array(sysm.interior, shape=(max(n, 1),)).copy()[:n]
            finally:
                L.lib.bn_nwk_sys_free(C.byref(sysm))
        finally:
            L.lib.bn_nwk_free(C.byref(w))


def read_network_case(base):
    return NetworkCase(base)


def write_pressures(path, p):
    """One %.16e per line (Matlab_amg/test_amg.m:68-72)."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    if L.lib.bn_write_pressures(str(path).encode(), L.dp(p), len(p)):
        raise OSError(L.last_error())


def write_labelled(path, nodes, p):
    """"node pressure" lines, 1-based nodes (Matlab_amg/label_solution.m:42-47)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    p = np.ascontiguousarray(p, dtype=np.float64)
    if len(nodes) != len(p):
        raise ValueError("one node per pressure")
    if L.lib.bn_write_labelled(str(path).encode(), L.ip(nodes), L.dp(p), len(p)):
        raise OSError(L.last_error())
