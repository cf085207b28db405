"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.c_double
_i = C.c_int
_i64 = C.c_int64
_p = C.c_void_p


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-B" if force else "-s", "liboracle.so"], cwd=_HERE, check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.or_elem_stiff.argtypes = [_d, _d, _d, _d, _p]
        L.or_elem_stiff_mr.argtypes = [_d, _d, _d, _d, _i, _p]
        L.or_create.restype = _p
        L.or_create.argtypes = [_i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _p, _i64]
        L.or_destroy.argtypes = [_p]
        L.or_nfree.restype = _i64
        L.or_nfree.argtypes = [_p, _i]
        L.or_ndof.restype = _i64
        L.or_ndof.argtypes = [_p, _i]
        L.or_assemble.argtypes = [_p, _p]
        L.or_assemble_mr.argtypes = [_p, _i, _p]
        L.or_setup_gmg.argtypes = [_p, _i, _d]
        L.or_refresh_gmg.argtypes = [_p]
        L.or_spmv.argtypes = [_p, _i, _p, _p]
        L.or_kelemprod.argtypes = [_p, _p, _p]
        L.or_smooth.argtypes = [_p, _i, _p, _p, _i, _p]
        L.or_vcycle.argtypes = [_p, _i, _p, _i, _i, _p]
        L.or_solve_mg.restype = _i
        L.or_solve_mg.argtypes = [_p, _p, _p, _i, _i, _d, _i, _p]
        L.or_wcycle.argtypes = [_p, _i, _p, _p, _i, _i, _p]
        L.or_vcycle_x0.argtypes = [_p, _i, _p, _p, _i, _i, _p]
        L.or_fmvcycle.argtypes = [_p, _i, _p, _i, _i, _p]
        L.or_solve_mg_cycle.restype = _i
        L.or_solve_mg_cycle.argtypes = [_p, _p, _p, _i, _i, _i, _d, _i, _p]
        L.or_csr_nnz.restype = _i64
        L.or_csr_nnz.argtypes = [_p, _i, _i]
        L.or_csr_export.argtypes = [_p, _i, _i, _p, _p, _p]
        L.or_level_free.argtypes = [_p, _i, _p]
        L.or_level_dims.argtypes = [_p, _i, _p]
        L.or_ngrids.restype = _i
        L.or_ngrids.argtypes = [_p]
        L.or_set_threads.restype = _i
        L.or_set_threads.argtypes = [_i]
        L.or_jacobi_diag.argtypes = [_p, _i, _p]
        L.or_filter_build.restype = _i64
        L.or_filter_build.argtypes = [_i, _i, _i, _d, _d, _d, _d, _i, _p, _p, _p, _p, _p]
        L.or_filter_apply.argtypes = [_i64, _p, _p, _p, _p, _p, _i, _p]
        L.or_grad0.argtypes = [_p, _p, _p, _p, _p]
        L.or_grad0_mr.argtypes = [_p, _i, _p, _p, _p]
        _LIB = L
    return _LIB


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's following calls; returns the previous count."""
    return int(lib().or_set_threads(int(n)))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def elem_stiff(el=1.0, eh=1.0, ew=1.0, nu=0.3) -> np.ndarray:
    k = np.zeros(576)
    lib().or_elem_stiff(el, eh, ew, nu, _ptr(k))
    return k.reshape(24, 24)


def elem_stiff_mr(n, el=1.0, eh=1.0, ew=1.0, nu=0.3) -> np.ndarray:
    k = np.zeros(576 * n ** 3)
    lib().or_elem_stiff_mr(el, eh, ew, nu, n, _ptr(k))
    return k.reshape(n ** 3, 24, 24)


class Oracle:
    """FP64 assembled-path oracle for one problem."""

    def __init__(self, prob, E0=1.0, p=3.0, Emin=1e-9):
        self.prob = prob
        fd = np.ascontiguousarray(prob.freedofs, dtype=np.int64)
        self.h = lib().or_create(prob.nelx, prob.nely, prob.nelz, prob.el, prob.eh, prob.ew,
                                 E0, prob.nu, p, Emin, _ptr(fd), fd.size)
        self.E0, self.p, self.Emin = E0, p, Emin

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def assemble(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert lib().or_assemble(self.h, _ptr(x)) == 0

    def assemble_mr(self, n, xd):
        xd = np.ascontiguousarray(xd, dtype=np.float64)
        assert lib().or_assemble_mr(self.h, n, _ptr(xd)) == 0

    def setup_gmg(self, ngrids, omega=0.6, allow_singular=False):
        rc = lib().or_setup_gmg(self.h, ngrids, omega)
        if rc == -2 and allow_singular:   # coarse Cholesky failed; P and A are built
            return
        if rc != 0:
            raise RuntimeError(f"or_setup_gmg failed ({rc})")

    def refresh(self):
        rc = lib().or_refresh_gmg(self.h)
        if rc != 0:
            raise RuntimeError(f"or_refresh_gmg failed ({rc})")

    @property
    def ngrids(self):
        return lib().or_ngrids(self.h)

    def nfree(self, level=0):
        return lib().or_nfree(self.h, level)

    def level_free(self, level):
        out = np.zeros(self.nfree(level), dtype=np.int64)
        lib().or_level_free(self.h, level, _ptr(out))
        return out

    def level_dims(self, level):
        out = np.zeros(3, dtype=np.int32)
        lib().or_level_dims(self.h, level, _ptr(out))
        return tuple(int(v) for v in out)

    def jacobi_diag(self, level):
        out = np.zeros(self.nfree(level))
        lib().or_jacobi_diag(self.h, level, _ptr(out))
        return out

    def spmv(self, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.nfree(level))
        lib().or_spmv(self.h, level, _ptr(x), _ptr(y))
        return y

    def kelemprod(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.nfree(0))
        lib().or_kelemprod(self.h, _ptr(x), _ptr(y))
        return y

    def smooth(self, level, b, x0, niter):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = np.zeros_like(b)
        lib().or_smooth(self.h, level, _ptr(b), _ptr(x0), niter, _ptr(x))
        return x

    def vcycle(self, b, nu1=1, nu2=1, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().or_vcycle(self.h, level, _ptr(b), nu1, nu2, _ptr(x))
        return x

    def wcycle(self, b, x0=None, nu1=1, nu2=1, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        lib().or_wcycle(self.h, level, _ptr(b), _ptr(x0a), nu1, nu2, _ptr(x))
        return x

    def vcycle_x0(self, b, x0, nu1=1, nu2=1, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        lib().or_vcycle_x0(self.h, level, _ptr(b), _ptr(x0a), nu1, nu2, _ptr(x))
        return x

    def fmvcycle(self, b, nu1=1, nu2=1, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().or_fmvcycle(self.h, level, _ptr(b), nu1, nu2, _ptr(x))
        return x

    def solve(self, f, x0=None, nu1=1, nu2=1, tol=1e-8, maxiter=5000, cycle=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.zeros_like(f) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        rr = np.zeros(1)
        it = lib().or_solve_mg_cycle(self.h, _ptr(x), _ptr(f), cycle, nu1, nu2, tol, maxiter, _ptr(rr))
        return x, int(it), float(rr[0])

    def csr(self, which, level):
        """which: 'A' or 'P'. Returns scipy.sparse.csr_matrix."""
        import scipy.sparse as sp
        w = 0 if which == "A" else 1
        nnz = lib().or_csr_nnz(self.h, w, level)
        n = self.nfree(level)
        m = self.nfree(level) if w == 0 else self.nfree(level + 1)
        rp = np.zeros(n + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().or_csr_export(self.h, w, level, _ptr(rp), _ptr(ci), _ptr(v))
        return sp.csr_matrix((v, ci, rp), shape=(n, m))

    def grad0(self, u_full, x, freedens=None):
        u_full = np.ascontiguousarray(u_full, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        fd = None if freedens is None else np.ascontiguousarray(freedens, dtype=np.uint8)
        g = np.zeros(self.prob.nelem)
        lib().or_grad0(self.h, _ptr(u_full), _ptr(x), _ptr(fd), _ptr(g))
        return g

    def grad0_mr(self, n, u_full, xd):
        u_full = np.ascontiguousarray(u_full, dtype=np.float64)
        xd = np.ascontiguousarray(xd, dtype=np.float64)
        g = np.zeros(xd.size)
        lib().or_grad0_mr(self.h, n, _ptr(u_full), _ptr(xd), _ptr(g))
        return g


def filter_matrix(nelx, nely, nelz, rmin, opfilter=2, el=1.0, eh=1.0, ew=1.0, fixed=None):
    """(W as scipy csr, w) following Implementation/ElemNeighborhood.m."""
    import scipy.sparse as sp
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
    L = lib()
    nnz = L.or_filter_build(nelx, nely, nelz, el, eh, ew, rmin, opfilter, _ptr(fx), None, None, None, None)
    n = nelx * nely * nelz
    rp = np.zeros(n + 1, dtype=np.int64)
    ci = np.zeros(nnz, dtype=np.int32)
    v = np.zeros(nnz)
    w = np.zeros(n)
    L.or_filter_build(nelx, nely, nelz, el, eh, ew, rmin, opfilter, _ptr(fx), _ptr(rp), _ptr(ci), _ptr(v), _ptr(w))
    return sp.csr_matrix((v, ci, rp), shape=(n, n)), w


def solve_reference(prob, x, ngrids, omega=0.6, tol=1e-8, maxiter=5000, E0=1.0, p=3.0, Emin=1e-9,
                    nu1=1, nu2=1):
    """Convenience: assemble, SetupGMG, SolveMG from x0 = 0. Returns (u_free, it, relres, oracle)."""
    o = Oracle(prob, E0=E0, p=p, Emin=Emin)
    o.assemble(x)
    o.setup_gmg(ngrids, omega)
    u, it, rr = o.solve(prob.f, nu1=nu1, nu2=nu2, tol=tol, maxiter=maxiter)
    return u, it, rr, o
