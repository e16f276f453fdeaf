"""Python mirror of the C_amg1 solver handle (C_amg1/src/amg_wrapper.c).

``AMGSolver`` follows the opaque-handle API one to one (create = build the
hierarchy + IC(0) and upload it to the GPU once; solve = PCG with the
AMG+IC preconditioner on the device), plus the extensions of
``include/bn_amg_ext.h`` (coarsening choice, FP32 V-cycle inside flexible
FP64 PCG, device-pointer solves, hierarchy introspection, adjoint, batch).
All numerics run in libbnamg.so; this module only moves arrays.
"""
import ctypes as C

import numpy as np

from . import _lib as L

__all__ = ["AMGSolver", "SolveReport"]


class SolveReport:
    """What one solve did (mirror of bn_solve_info)."""

    def __init__(self, info):
        self.iterations = info.iterations
        self.rel_bnorm = info.rel_bnorm
        self.relres = info.relres
        self.solve_ms = info.solve_ms
        self.wall_ms = info.wall_ms
        self.setup_ms = info.setup_ms
        self.upload_ms = info.upload_ms
        self.num_levels = info.num_levels

    def __repr__(self):
        return (f"SolveReport(iterations={self.iterations}, rel_bnorm={self.rel_bnorm:.3e}, "
                f"relres={self.relres:.3e}, solve_ms={self.solve_ms:.3f})")


def _csr(ia, ja, val):
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return ia, ja, val


class AMGSolver:
    """AMG(+IC)-preconditioned PCG solver on the GPU.

    ``AMGSolver(ia, ja, val, threshold, min_size, max_levels, nu)`` is
    ``amg_solver_create`` (PMIS, CF-GS, symmetrised input).  Keyword
    arguments select the extensions: ``coarsening`` ("rs" | "pmis" |
    "hmis"), ``precision`` ("fp64" | "fp32"), ``coarse`` ("auto" | "dense"
    | "icpcg"), ``fine_ic`` ("exact" | "jacobi"), ``flexible``, ``stop``
    ("bnorm" | "relres"), ``smoother``, ``block_smoother`` (chunked l1 block
    Gauss-Seidel on every level, one launch per sweep) with ``chunk_rows``,
    ``strength`` ("c": C_amg1's rule | "mex": the MATLAB MEX rule, which
    clamps -a_ij at 0), ``gs_exact`` (SMOOTHER_CF_GS sweeps in the reference's
    exact cf_order, level-scheduled, instead of the C-then-F multicolour
    sweep; SMOOTHER_GS is always the exact natural-order sweep).
    """

    _COARSEN = {"rs": L.COARSEN_RS, "pmis": L.COARSEN_PMIS, "hmis": L.COARSEN_HMIS}
    _COARSE = {"auto": L.COARSE_AUTO, "dense": L.COARSE_DENSE, "icpcg": L.COARSE_ICPCG}
    _IC = {"exact": L.IC_EXACT, "jacobi": L.IC_JACOBI, "none": L.IC_NONE}
    _SMOOTH = {"l1": L.SMOOTHER_L1, "gs": L.SMOOTHER_GS, "cf_gs": L.SMOOTHER_CF_GS,
               "jacobi": L.SMOOTHER_JACOBI}

    def __init__(self, ia, ja, val, threshold=0.25, min_size=4096, max_levels=20, nu=1, *,
                 coarsening=None, precision="fp64", coarse="auto", fine_ic="exact",
                 ic_sweeps=3, flexible=None, stop="bnorm", smoother="cf_gs", max_colours=32,
                 symmetrize=True, verbose=False, graph=True, block_smoother=False, chunk_rows=0,
                 strength="c", gs_exact=False):
        self.ia, self.ja, self.val = _csr(ia, ja, val)
        self.n = len(self.ia) - 1
        plain = (coarsening is None and precision == "fp64" and coarse == "auto"
                 and fine_ic == "exact" and stop == "bnorm" and smoother == "cf_gs"
                 and flexible is None and symmetrize and not verbose and max_colours == 32
                 and graph and not block_smoother and strength == "c" and not gs_exact)
        if plain:
            h = L.lib.amg_solver_create(self.n, len(self.ja), L.ip(self.ia), L.ip(self.ja),
                                        L.dp(self.val), threshold, min_size, max_levels, nu)
        else:
            o = L.Options()
            L.lib.bn_options_default(C.byref(o))
            o.coarsening = self._COARSEN[coarsening or "pmis"]
            o.threshold = threshold
            o.min_size = min_size
            o.max_levels = max_levels
            o.nu = nu
            o.smoother = self._SMOOTH[smoother]
            o.precision = L.PREC_FP32 if precision == "fp32" else L.PREC_FP64
            o.coarse_solver = self._COARSE[coarse]
            o.fine_ic = self._IC[fine_ic]
            o.ic_sweeps = ic_sweeps
            o.max_colours = max_colours
            if flexible is None:
                flexible = precision == "fp32" or fine_ic == "jacobi"
            o.flexible = int(bool(flexible))
            o.stop = L.STOP_RELRES if stop == "relres" else L.STOP_BNORM
            o.symmetrize = int(bool(symmetrize))
            o.verbose = int(bool(verbose))
            o.graph = int(bool(graph))
            o.block_smoother = int(bool(block_smoother))
            if chunk_rows:
                o.chunk_rows = int(chunk_rows)
            o.strength = {"c": L.STRENGTH_C, "mex": L.STRENGTH_MEX}[strength]
            o.gs_exact = int(bool(gs_exact))
            h = L.lib.bn_solver_create(self.n, len(self.ja), L.ip(self.ia), L.ip(self.ja),
                                       L.dp(self.val), C.byref(o))
        if not h:
            raise ValueError(f"solver creation failed: {L.last_error()}")
        self.h = h

    # -- reference API (amg_wrapper.c) ---------------------------------
    def solve(self, b, x=None, tol=1e-8, maxiter=1000, print_level=0):
        """PCG solve; returns (x, iterations).  x is the initial guess."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x is None else np.array(x, dtype=np.float64, copy=True)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_info(self.h, L.dp(b), L.dp(x), tol, maxiter, print_level,
                                        C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return x, it

    def solve_device(self, b_ptr, x_ptr, tol=1e-8, maxiter=1000, print_level=0):
        """Solve with device pointers (e.g. torch ``data_ptr()``)."""
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_device(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), tol,
                                          maxiter, print_level, C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return it

    def update(self, ia, ja, val):
        ia, ja, val = _csr(ia, ja, val)
        self._check(L.lib.amg_solver_update(self.h, len(ia) - 1, len(ja), L.ip(ia), L.ip(ja), L.dp(val)))
        self.ia, self.ja, self.val = ia, ja, val

    def set_matrix(self, ia, ja, val):
        ia, ja, val = _csr(ia, ja, val)
        self._check(L.lib.amg_solver_set_matrix(self.h, len(ia) - 1, len(ja), L.ip(ia), L.ip(ja),
                                                L.dp(val)))

    def set_smoother(self, smoother, omega=0.5):
        s = self._SMOOTH[smoother] if isinstance(smoother, str) else int(smoother)
        L.lib.amg_solver_set_smoother(self.h, s, omega)

    def set_mode(self, precision="fp64", fine_ic="exact", ic_sweeps=0, flexible=None, stop="bnorm"):
        if flexible is None:
            flexible = precision == "fp32" or fine_ic == "jacobi"
        L.lib.bn_solver_set_mode(self.h, L.PREC_FP32 if precision == "fp32" else L.PREC_FP64,
                                 self._IC[fine_ic], ic_sweeps,
                                 int(bool(flexible)), L.STOP_RELRES if stop == "relres" else L.STOP_BNORM)

    def size(self):
        return L.lib.amg_solver_size(self.h)

    def num_levels(self):
        return L.lib.amg_solver_num_levels(self.h)

    def free(self):
        if getattr(self, "h", None):
            L.lib.amg_solver_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- extensions ------------------------------------------------------
    def precondition(self, g, which="ichol"):
        g = np.ascontiguousarray(g, dtype=np.float64)
        z = np.zeros_like(g)
        self._check(L.lib.bn_solver_precond(self.h, L.dp(g), L.dp(z), 0 if which == "ichol" else 1))
        return z

    def adjoint_grad(self, p, dLdp, eu, ev, g, pdir, tol=1e-10, maxiter=1000):
        """lambda = A^{-1} dL/dp, dL/dg_e and dL/dlog r_e (host arrays)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        dLdp = np.ascontiguousarray(dLdp, dtype=np.float64)
        eu = np.ascontiguousarray(eu, dtype=np.int32)
        ev = np.ascontiguousarray(ev, dtype=np.int32)
        g = np.ascontiguousarray(g, dtype=np.float64)
        pdir = np.ascontiguousarray(pdir, dtype=np.float64)
        lam = np.zeros(self.n)
        dg = np.zeros(len(eu))
        dlr = np.zeros(len(eu))
        it = L.lib.bn_solver_adjoint_grad(self.h, L.dp(p), L.dp(dLdp), len(eu), L.ip(eu), L.ip(ev),
                                          L.dp(g), L.dp(pdir), tol, maxiter, L.dp(lam), L.dp(dg),
                                          L.dp(dlr))
        self._check(it)
        return lam, dg, dlr, it

    def adjoint_grad_noalloc(self, p, dLdp, eu, ev, g, pdir, tol=1e-10, maxiter=1000):
        """adjoint_grad without device allocation (static scratch; host
        arrays, same results)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        dLdp = np.ascontiguousarray(dLdp, dtype=np.float64)
        eu = np.ascontiguousarray(eu, dtype=np.int32)
        ev = np.ascontiguousarray(ev, dtype=np.int32)
        g = np.ascontiguousarray(g, dtype=np.float64)
        pdir = np.ascontiguousarray(pdir, dtype=np.float64)
        lam = np.zeros(self.n)
        dg = np.zeros(len(eu))
        dlr = np.zeros(len(eu))
        it = L.lib.bn_solver_adjoint_grad_noalloc(self.h, L.dp(p), L.dp(dLdp), len(eu), L.ip(eu), L.ip(ev),
                                                  L.dp(g), L.dp(pdir), tol, maxiter, L.dp(lam), L.dp(dg),
                                                  L.dp(dlr))
        self._check(it)
        return lam, dg, dlr, it

    def adjoint_grad_device(self, p_ptr, dLdp_ptr, m, eu_ptr, ev_ptr, g_ptr, pdir_ptr, nd, lam_ptr,
                            dLdg_ptr=None, dLdlogr_ptr=None, tol=1e-10, maxiter=1000):
        """The adjoint gradient on device pointers; lam_ptr is the initial
        guess on entry.  Returns the adjoint iteration count."""
        it = L.lib.bn_solver_adjoint_grad_device(self.h, p_ptr, dLdp_ptr, m, eu_ptr, ev_ptr, g_ptr, pdir_ptr,
                                                 nd, tol, maxiter, lam_ptr, dLdg_ptr, dLdlogr_ptr)
        self._check(it)
        info = L.SolveInfo()
        L.lib.bn_solver_last_info(self.h, C.byref(info))
        self.report = SolveReport(info)
        return it

    def set_values_device(self, nnz, val_ptr):
        """New solve-matrix values (device pointer, nnz entries in the
        creation order); the preconditioner is kept."""
        self._check(L.lib.bn_solver_set_values_device(self.h, nnz, val_ptr))

    def rebuild_values_device(self, nnz, val_ptr):
        """New solve-matrix values (device pointer, creation entry order) and
        a value-only rebuild of the whole hierarchy on the device
        (amg_solver_update's result, computed on the GPU)."""
        self._check(L.lib.bn_solver_rebuild_values_device(self.h, nnz, val_ptr))

    def coarse_inverse(self):
        """The solver's dense inverse of the coarsest operator (n_c x n_c)."""
        n = C.c_int(0)
        self._check(L.lib.bn_solver_coarse_inverse(self.h, C.byref(n), None))
        out = np.zeros((n.value, n.value))
        self._check(L.lib.bn_solver_coarse_inverse(self.h, C.byref(n), out.ctypes.data_as(L.DP)))
        return out

    def solve_batch(self, B, tol=1e-8, maxiter=1000):
        """B: (k, n) right-hand sides; returns X (k, n) and per-column counts."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        k = B.shape[0]
        X = np.zeros_like(B)
        its = np.zeros(k, np.int32)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_batch(self.h, k, L.dp(B), L.dp(X), tol, maxiter, L.ip(its),
                                         C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return X, its

    def solve_segments(self, b, seg_ptr, x0=None, tol=1e-8, maxiter=1000):
        """Independent systems stacked block-diagonally (rows of system g:
        seg_ptr[g] .. seg_ptr[g+1]-1) solved in one PCG with per-system
        scalars (bn_solver_solve_segments); returns x and per-system
        iteration counts."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        seg = np.ascontiguousarray(seg_ptr, dtype=np.int32)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
        its = np.zeros(len(seg) - 1, np.int32)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_segments(self.h, len(seg) - 1, L.ip(seg), L.dp(b), L.dp(x), tol, maxiter,
                                            L.ip(its), C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return x, its

    def solve_segments_device(self, seg_ptr, b_ptr, x_ptr, tol=1e-8, maxiter=1000):
        seg = np.ascontiguousarray(seg_ptr, dtype=np.int32)
        its = np.zeros(len(seg) - 1, np.int32)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_segments_device(self.h, len(seg) - 1, L.ip(seg), C.c_void_p(b_ptr),
                                                   C.c_void_p(x_ptr), tol, maxiter, L.ip(its), C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return its

    def solve_block(self, B, X0=None, tol=1e-8, maxiter=1000):
        """Block PCG over the k <= 32 right-hand sides B (k, n) in one
        solve (bn_solver_solve_block); returns X (k, n) and per-column
        iteration counts."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        k = B.shape[0]
        X = np.zeros_like(B) if X0 is None else np.array(X0, dtype=np.float64, order="C")
        its = np.zeros(k, np.int32)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_block(self.h, k, L.dp(B), L.dp(X), tol, maxiter, L.ip(its), C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return X, its

    def solve_block_device(self, k, b_ptr, x_ptr, tol=1e-8, maxiter=1000):
        """Device-pointer block solve: b_ptr, x_ptr hold k x n doubles."""
        its = np.zeros(k, np.int32)
        info = L.SolveInfo()
        it = L.lib.bn_solver_solve_block_device(self.h, k, C.c_void_p(b_ptr), C.c_void_p(x_ptr), tol, maxiter,
                                                L.ip(its), C.byref(info))
        self._check(it)
        self.report = SolveReport(info)
        return its

    def kernel_census(self, b_ptr, x_ptr, tol=1e-8, maxiter=1000, reps=3, max_symbols=64):
        """One solve (device pointers, x the initial guess) with every
        bandwidth-kernel launch timed by HIP events, then each launch shape
        timed cold; returns a list of dicts per kernel symbol, slowest first:
        name, launches, shapes, bytes, us_solve, us_cold (per solve)."""
        out = (L.KernelStat * max_symbols)()
        n = C.c_int(0)
        self._check(L.lib.bn_solver_kernel_census(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), tol, maxiter, reps,
                                                  out, max_symbols, C.byref(n)))
        res = [{"name": o.name.decode(), "launches": o.launches, "shapes": o.shapes, "bytes": o.bytes,
                "us_solve": o.us_solve, "us_cold": o.us_cold} for o in out[:n.value]]
        return sorted(res, key=lambda r: -r["us_solve"])

    def level_dims(self, lev):
        d = np.zeros(8, np.int32)
        self._check(L.lib.bn_solver_level_dims(self.h, lev, L.ip(d)))
        return d

    def level(self, lev):
        """Host copies of level `lev`: dict with A, P, S (ia, ja, val),
        mis, cf_order, l1 and, on the coarsest level, Lc."""
        n, nnzA, nc, nnzP, nnzS, nnzLc, nnzLf, has = self.level_dims(lev)
        out = {}

        def grab(what, rows, nnz):
            ia = np.zeros(rows + 1, np.int32)
            ja = np.zeros(nnz, np.int32)
            v = np.zeros(nnz)
            self._check(L.lib.bn_solver_level_csr(self.h, lev, what, L.ip(ia), L.ip(ja), L.dp(v)))
            return ia, ja, v

        out["A"] = grab(0, n, nnzA)
        if has:
            out["P"] = grab(1, n, nnzP)
            out["S"] = grab(2, n, nnzS)
            mis = np.zeros(n, np.int32)
            co = np.zeros(n, np.int32)
            l1 = np.zeros(n)
            self._check(L.lib.bn_solver_level_split(self.h, lev, L.ip(mis), L.ip(co), L.dp(l1)))
            out["mis"], out["cf_order"], out["l1"] = mis, co, l1
        if nnzLc:
            out["Lc"] = self.ichol(1, n, nnzLc)
        return out

    def ichol(self, which=0, n=None, nnz=None):
        if n is None:
            n = self.n
            nnz = self.level_dims(0)[6]
        ia = np.zeros(n + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        self._check(L.lib.bn_solver_ichol(self.h, which, L.ip(ia), L.ip(ja), L.dp(v)))
        return ia, ja, v

    def schedule(self, lev):
        nc, gs, lanes = C.c_int(), C.c_int(), C.c_int()
        self._check(L.lib.bn_solver_schedule(self.h, lev, C.byref(nc), C.byref(gs), C.byref(lanes)))
        return nc.value, bool(gs.value), lanes.value

    def ic_levels(self):
        f, b = C.c_int(), C.c_int()
        self._check(L.lib.bn_solver_ic_levels(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    @staticmethod
    def _check(rc):
        if rc in (L.BN_ERR_NO_DEVICE, L.BN_ERR_HIP, L.BN_ERR_ARG):
            raise RuntimeError(f"libbnamg error {rc}: {L.last_error()}")
        return rc
