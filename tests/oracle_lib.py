"""ctypes access to the oracle (oracle/liboracle.so) and, when it was
built, the compiled reference (oracle/_ref/libamgref.so).

Test infrastructure only: both libraries expose the same flat entry points
(oracle/harness.c) over an amg_data hierarchy.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libamgref.so")

IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)


def ip(a):
    return a.ctypes.data_as(IP)


def dp(a):
    return a.ctypes.data_as(DP)


COARSEN = {"rs": 0, "pmis": 1, "hmis": 2}


class Harness:
    """One hierarchy implementation behind oracle/harness.c."""

    def __init__(self, path):
        self.path = path
        self.lib = l = C.CDLL(path)
        l.h_build.restype = C.c_void_p
        l.h_build.argtypes = [C.c_int, C.c_int, IP, IP, DP, C.c_double, C.c_int, C.c_int, C.c_int,
                              C.c_int, C.c_int, C.c_int]
        l.h_update.argtypes = [C.c_void_p, DP, C.c_int]
        l.h_free.argtypes = [C.c_void_p]
        l.h_num_levels.argtypes = [C.c_void_p]
        l.h_level_dims.argtypes = [C.c_void_p, C.c_int, IP]
        l.h_level_csr.argtypes = [C.c_void_p, C.c_int, C.c_int, IP, IP, DP]
        l.h_level_split.argtypes = [C.c_void_p, C.c_int, IP, IP, DP]
        l.h_ichol.argtypes = [C.c_void_p, C.c_int, IP, IP, DP]
        l.h_set_smoother.argtypes = [C.c_void_p, C.c_int, C.c_double]
        l.h_solve.argtypes = [C.c_void_p, DP, DP, C.c_double, C.c_int, C.c_int, C.c_int, DP]
        l.h_precond.argtypes = [C.c_void_p, DP, DP]
        l.h_vcycle.argtypes = [C.c_void_p, DP, DP]
        l.h_backslash.argtypes = [C.c_void_p, C.c_int, DP, DP]
        l.h_fwdslash.argtypes = [C.c_void_p, C.c_int, DP, DP]
        l.h_ichol_solve.argtypes = [C.c_void_p, C.c_int, DP, DP]
        l.h_fine_csr.argtypes = [C.c_void_p, IP, IP, DP]
        l.h_fine_nnz.argtypes = [C.c_void_p]
        # raw restated routines (same names and types as C_amg1)
        for name in ("abyvg", "abyvcs"):
            getattr(l, name).argtypes = [IP, IP, DP, DP, C.c_int, DP]
        l.abyvam.argtypes = [DP, IP, IP, DP, DP, C.c_int, DP]
        l.vbya.argtypes = [IP, IP, DP, DP, C.c_int, C.c_int, DP]

    def build(self, ia, ja, val, theta=0.25, min_size=4096, max_levels=20, nu=1, coarsening="pmis",
              smoother=2, symmetrize=True):
        self._keep = (np.ascontiguousarray(ia, np.int32), np.ascontiguousarray(ja, np.int32),
                      np.ascontiguousarray(val, np.float64))
        ia, ja, val = self._keep
        c = COARSEN[coarsening] if isinstance(coarsening, str) else coarsening
        return Hier(self, self.lib.h_build(len(ia) - 1, len(ja), ip(ia), ip(ja), dp(val), theta,
                                           min_size, max_levels, nu, c, smoother, int(symmetrize)))


class Hier:
    def __init__(self, hx, h):
        self.hx, self.h, self.lib = hx, h, hx.lib

    def __del__(self):
        try:
            self.lib.h_free(self.h)
        except Exception:
            pass

    def num_levels(self):
        return self.lib.h_num_levels(self.h)

    def dims(self, lev):
        d = np.zeros(8, np.int32)
        self.lib.h_level_dims(self.h, lev, ip(d))
        return d

    def level(self, lev):
        n, nnzA, nc, nnzP, nnzS, nnzLc, nnzLf, has = self.dims(lev)
        out = {}

        def grab(what, nnz):
            ia = np.zeros(n + 1, np.int32)
            ja = np.zeros(nnz, np.int32)
            v = np.zeros(nnz)
            self.lib.h_level_csr(self.h, lev, what, ip(ia), ip(ja), dp(v))
            return ia, ja, v

        out["A"] = grab(0, nnzA)
        if has:
            out["P"] = grab(1, nnzP)
            out["S"] = grab(2, nnzS)
            mis = np.zeros(n, np.int32)
            co = np.zeros(n, np.int32)
            l1 = np.zeros(n)
            self.lib.h_level_split(self.h, lev, ip(mis), ip(co), dp(l1))
            out["mis"], out["cf_order"], out["l1"] = mis, co, l1
        if nnzLc:
            ia = np.zeros(n + 1, np.int32)
            ja = np.zeros(nnzLc, np.int32)
            v = np.zeros(nnzLc)
            self.lib.h_ichol(self.h, 1, ip(ia), ip(ja), dp(v))
            out["Lc"] = (ia, ja, v)
        return out

    def fine_ichol(self):
        n = self.dims(0)[0]
        nnz = self.dims(0)[6]
        ia = np.zeros(n + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        self.lib.h_ichol(self.h, 0, ip(ia), ip(ja), dp(v))
        return ia, ja, v

    def fine_csr(self):
        n = self.dims(0)[0]
        nnz = self.lib.h_fine_nnz(self.h)
        ia = np.zeros(n + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        self.lib.h_fine_csr(self.h, ip(ia), ip(ja), dp(v))
        return ia, ja, v

    def solve(self, b, tol, maxiter, precond=0, x0=None, print_level=0):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True)
        sec = C.c_double()
        it = self.lib.h_solve(self.h, dp(b), dp(x), tol, maxiter, print_level, precond, C.byref(sec))
        return x, it, sec.value

    def precond(self, g):
        g = np.ascontiguousarray(g, np.float64)
        z = np.zeros_like(g)
        self.lib.h_precond(self.h, dp(g), dp(z))
        return z

    def vcycle(self, g):
        g = np.ascontiguousarray(g, np.float64)
        z = np.zeros_like(g)
        self.lib.h_vcycle(self.h, dp(g), dp(z))
        return z

    def set_smoother(self, smoother, omega=0.5):
        self.lib.h_set_smoother(self.h, smoother, omega)

    def update(self, val, symmetrize=True):
        val = np.ascontiguousarray(val, np.float64)
        self.lib.h_update(self.h, dp(val), int(symmetrize))


_cache = {}


def oracle():
    if "o" not in _cache:
        _cache["o"] = Harness(ORACLE_PATH)
    return _cache["o"]


def reference():
    """The compiled reference, or None when oracle/_ref was not built."""
    if "r" not in _cache:
        _cache["r"] = Harness(REF_PATH) if os.path.exists(REF_PATH) else None
    return _cache["r"]
