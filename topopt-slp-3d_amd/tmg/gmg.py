"""Host-side mirror of the reference's MATLAB call surface for the equilibrium
solve, backed by libtmg (hand-written gfx950 HIP kernels).

Functions keep the reference's names and positional arguments so that the
``Implementation/StrSLP.m`` call sequence reads the same:

    K = GlobalStiffMatrix(str, kv, iK, jK, x, p, elem, emin, mr, strDens)
    [A, P, M, L, U, R, perm] = SetupGMG(K, str, mg, elem)
    [u(freedofs), it] = SolveMG(u0, str.f, A, P, M, L, U, R, perm, ngrids, cycle,
                                smoother, smoothiter1, smoothiter2, tolMG, maxiterMG,
                                opchol, opcond, pcgp)
    Fgrad = Grad(str, u, p, emin, k, x, Dofs, W, w, opfilter, mr, strDens)

The stiffness matrix is never assembled: ``GlobalStiffMatrix`` returns a
description of K(x) and ``SetupGMG`` builds the device-resident Galerkin
hierarchy from it. ``A`` is the hierarchy handle; ``P`` is a list of
device-backed prolongation operators; ``M, L, U, R, perm`` are kept for
argument-position compatibility and refer to the same hierarchy.

Non-convergence is silent, as in the reference (Implementation/SolveMG.m:30);
invalid arguments, unsupported options and device errors raise TmgError.
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import TMG_FP64, TMG_MIXED, TmgError, check, lib

_PREC = {"fp64": TMG_FP64, "mixed": TMG_MIXED, TMG_FP64: TMG_FP64, TMG_MIXED: TMG_MIXED}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _get(o, name, default=None):
    if isinstance(o, dict):
        return o.get(name, default)
    return getattr(o, name, default)


class Solver:
    """Device-resident GMG-PCG hierarchy for one structured-grid problem."""

    def __init__(self, prob, ngrids, precision="fp64", omega=0.6, smoothiter1=1, smoothiter2=1,
                 tol=1e-8, maxiter=5000, p=3.0, emin=1e-9, E0=None, mr=1, cycle=0):
        self.prob = prob
        self.cycle = int(cycle)
        self.mr = int(mr)
        self.ndens = prob.nelem * self.mr ** 3   # densities / gradient entries (density mesh)
        self.precision = _PREC[precision]
        fd = np.ascontiguousarray(prob.freedofs, dtype=np.int64)
        self._fd = fd
        pr = _lib.Problem_t(prob.nelx, prob.nely, prob.nelz, prob.el, prob.eh, prob.ew,
                            float(prob.E if E0 is None else E0), prob.nu, float(p), float(emin),
                            fd.ctypes.data_as(C.POINTER(C.c_int64)), fd.size)
        mg = _lib.MgParams_t(ngrids, self.cycle, 0, omega, smoothiter1, smoothiter2, tol, maxiter, 0, 60000)
        h = C.c_void_p()
        if self.mr == 1:
            rc = lib().tmg_setup(C.byref(pr), C.byref(mg), self.precision, C.byref(h))
        else:
            rc = lib().tmg_setup_mr(C.byref(pr), C.byref(mg), self.precision, self.mr, C.byref(h))
        if rc != 0:
            raise TmgError(rc, _lib.last_error(None))
        self.h = h
        self.ngrids = ngrids
        self.nfree = fd.size
        self.params = (smoothiter1, smoothiter2, float(tol), maxiter)

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            lib().tmg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, allow_nc=False):
        return check(rc, self.h, allow_nc)

    # -- operators ---------------------------------------------------------
    def update_density(self, x):
        """Refresh K(x), the Galerkin hierarchy, Jacobi diagonals and coarse factor."""
        if hasattr(x, "data_ptr"):  # device tensor (float64, n_el)
            return self._chk(lib().tmg_update_density_device(self.h, C.c_void_p(x.data_ptr())))
        x = _f64(x)
        if x.size != self.ndens:
            raise ValueError("density vector has the wrong length")
        return self._chk(lib().tmg_update_density(self.h, _dp(x)))

    def level_info(self, level):
        nel = (C.c_int * 3)()
        nf = C.c_int64()
        self._chk(lib().tmg_level_info(self.h, level, nel, C.byref(nf), None))
        fd = np.zeros(nf.value, dtype=np.int64)
        self._chk(lib().tmg_level_info(self.h, level, nel, C.byref(nf), fd.ctypes.data_as(C.POINTER(C.c_int64))))
        return tuple(nel), fd

    def nfree_level(self, level):
        nel = (C.c_int * 3)()
        nf = C.c_int64()
        self._chk(lib().tmg_level_info(self.h, level, nel, C.byref(nf), None))
        return nf.value

    def solve(self, f, x0=None):
        """Host free-DOF vectors. Returns (x, it, relres) with it = iterations + 1."""
        f = _f64(f)
        x0 = np.zeros_like(f) if x0 is None else _f64(x0)
        x = np.zeros_like(f)
        it = C.c_int()
        rr = C.c_double()
        rc = self._chk(lib().tmg_solve(self.h, _dp(x0), _dp(f), _dp(x), C.byref(it), C.byref(rr)), True)
        return x, it.value, rr.value, rc

    def solve_device(self, x, f):
        """Device tensors (float64, full length 3*nnodes, fixed DOFs 0); x is updated in place.
        The solver runs on its own HIP stream: pending work of the caller's
        stream on x and f is waited for first."""
        if x.dtype != f.dtype or str(x.dtype) != "torch.float64" or x.numel() != self.prob.ndof or f.numel() != self.prob.ndof:
            raise ValueError("solve_device needs float64 tensors of length 3*nnodes")
        if not (x.is_contiguous() and f.is_contiguous()):
            raise ValueError("solve_device needs contiguous tensors")
        import torch
        torch.cuda.current_stream(x.device).synchronize()
        it = C.c_int()
        rr = C.c_double()
        rc = self._chk(lib().tmg_solve_device(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(f.data_ptr()),
                                              C.byref(it), C.byref(rr)), True)
        return it.value, rr.value, rc

    def vcycle(self, b):
        b = _f64(b)
        x = np.zeros_like(b)
        self._chk(lib().tmg_vcycle(self.h, _dp(b), _dp(x)))
        return x

    def apply(self, level, x):
        x = _f64(x)
        y = np.zeros_like(x)
        self._chk(lib().tmg_apply(self.h, level, _dp(x), _dp(y)))
        return y

    def smooth(self, level, b, x0, niter):
        b, x0 = _f64(b), _f64(x0)
        x = np.zeros_like(b)
        self._chk(lib().tmg_smooth(self.h, level, _dp(b), _dp(x0), niter, _dp(x)))
        return x

    def prolong(self, level, xc):
        xc = _f64(xc)
        xf = np.zeros(self.nfree_level(level))
        self._chk(lib().tmg_prolong(self.h, level, _dp(xc), _dp(xf)))
        return xf

    def restrict(self, level, xf):
        xf = _f64(xf)
        xc = np.zeros(self.nfree_level(level + 1))
        self._chk(lib().tmg_restrict(self.h, level, _dp(xf), _dp(xc)))
        return xc

    def solve_stats(self):
        """(iteration at which the last solve switched to the FP64 preconditioner or 0,
        number of switches so far) -- mixed precision only."""
        a, b = C.c_int(), C.c_int()
        self._chk(lib().tmg_solve_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def device_bytes(self):
        """Device bytes the handle holds (allocations rounded up to 256 bytes)."""
        return int(lib().tmg_device_bytes(self.h))

    def time_kernel(self, which, nrep=20):
        """Average launch time (ms) of a solver kernel, HIP events on the handle's stream."""
        ms = C.c_double()
        self._chk(lib().tmg_time_kernel(self.h, which, nrep, C.byref(ms)))
        return ms.value

    def time_in_solve(self, which):
        """Average launch time (ms) of a solver kernel (0 FP64 apply, 8 fine
        post-sweep) inside one solve on the last right-hand side, and the
        number of launches averaged; overwrites the handle's solution."""
        ms, n = C.c_double(), C.c_int()
        self._chk(lib().tmg_time_in_solve(self.h, which, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def grad(self, u_full, x, freedens=None, filt=None):
        u_full, x = _f64(u_full), _f64(x)
        g = np.zeros(self.ndens)
        fd = None
        if freedens is not None:
            fdm = np.ascontiguousarray(freedens, dtype=np.uint8)
            fd = fdm.ctypes.data_as(C.POINTER(C.c_uint8))
        self._chk(lib().tmg_grad(self.h, _dp(u_full), _dp(x), fd, filt.h if filt is not None else None, _dp(g)))
        return g


class Filter:
    """Device density filter [W, w] = ElemNeighborhood(str, rmin, opfilter)."""

    def __init__(self, nelx, nely, nelz, rmin, opfilter=2, el=1.0, eh=1.0, ew=1.0, fixed=None):
        h = C.c_void_p()
        fx = None
        if fixed is not None:
            self._fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
            fx = self._fixed.ctypes.data_as(C.POINTER(C.c_uint8))
        rc = lib().tmg_filter_create(nelx, nely, nelz, el, eh, ew, rmin, opfilter, fx, C.byref(h))
        if rc != 0:
            raise TmgError(rc, _lib.last_error(None))
        self.h = h
        self.n = nelx * nely * nelz

    def apply(self, x, trans=False):
        x = _f64(x)
        y = np.zeros(self.n)
        rc = lib().tmg_filter_apply(self.h, _dp(x), 1 if trans else 0, _dp(y))
        if rc:
            raise TmgError(rc, "filter apply failed")
        return y

    def rowsum(self):
        w = np.zeros(self.n)
        rc = lib().tmg_filter_rowsum(self.h, _dp(w))
        if rc:
            raise TmgError(rc, "filter rowsum failed")
        return w

    def close(self):
        if getattr(self, "h", None):
            lib().tmg_filter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# MATLAB call surface
# ---------------------------------------------------------------------------
class _Prolongation:
    """Device-backed P{level}: ``P @ xc`` prolongs, ``P.T @ xf`` restricts."""

    def __init__(self, solver, level, transpose=False):
        self.s, self.level, self.t = solver, level, transpose

    @property
    def T(self):
        return _Prolongation(self.s, self.level, not self.t)

    def __matmul__(self, v):
        return self.s.restrict(self.level, v) if self.t else self.s.prolong(self.level, v)


def _mr_factor(mr):
    """mr.n of a multiresolution struct (1 without multiresolution); only
    mr.n = mr.d = 2 with mr.interp = false is supported."""
    if mr is None or not _get(mr, "op", False):
        return 1
    n, d = int(_get(mr, "n", 2)), int(_get(mr, "d", _get(mr, "n", 2)))
    if n != 2 or d != 2 or _get(mr, "interp", False):
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "multiresolution is supported with mr.n = mr.d = 2, mr.interp = false")
    return 2


def GlobalStiffMatrix(str, kv, iK, jK, x, p, elem, emin, mr=None, strDens=None):
    """K = sum_e (emin + (E - emin) x_e^p) Ke restricted to freedofs
    (Implementation/GlobalStiffMatrix.m:16-18,64-65); with mr.op, each element
    sums its 8 density sub-elements' integrands, x on the density mesh
    (Implementation/GlobalStiffMatrix.m:19-62). Returns a description of K(x);
    the matrix is applied element by element on the device."""
    if _get(elem, "deg", 1) != 1:
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "only Q1 elements (elem.deg = 1) are supported")
    n = _mr_factor(mr)
    x = np.asarray(x, dtype=np.float64)
    if x.size != str.nelem * n ** 3:
        raise TmgError(_lib.TMG_ERR_ARG, "x does not match the element (or density) mesh")
    return SimpleNamespace(str=str, x=x, p=float(p), emin=float(emin), mr=n)


def SetupGMG(K, str, mg, elem, precision=None):
    """[A,P,M,L,U,R,perm] = SetupGMG(K,str,mg,elem) (Implementation/SetupGMG.m:1-78)."""
    if _get(elem, "deg", 1) != 1 or _get(elem, "type", 1) != 1:
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "only Q1 Lagrange elements are supported")
    prec = precision if precision is not None else _get(mg, "precision", "fp64")
    s = Solver(str, int(_get(mg, "ngrids")), precision=prec, omega=float(_get(mg, "omega", 0.6)),
               smoothiter1=int(_get(mg, "smoothiter1", 1)), smoothiter2=int(_get(mg, "smoothiter2", 1)),
               tol=float(_get(mg, "tolMG", 1e-8)), maxiter=int(_get(mg, "maxiterMG", 5000)),
               p=K.p, emin=K.emin, cycle=int(_get(mg, "cycle", 0)), mr=getattr(K, "mr", 1))
    if int(_get(mg, "smoother", 0)) != 0 or int(_get(mg, "opchol", 0)) != 0:
        s.close()
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "only the Jacobi smoother and coarse Cholesky are supported")
    s.update_density(K.x)
    P = [_Prolongation(s, i) for i in range(1, s.ngrids)]
    return s, P, s, s, s, s, s


def RefreshGMG(A, K):
    """Galerkin refresh after a new K (Implementation/StrSLP.m:319,345-365)."""
    A.update_density(K.x)
    return A


def SolveMG(x, f, Ac, P, M, L, U, R, perm, ngrids, cycle, smoother, smoothiter1, smoothiter2,
            tolMG, maxiterMG, opchol, opcond, pcgp=None):
    """[x,it] = SolveMG(...) (Implementation/SolveMG.m:1-55). The hierarchy's
    smoothing counts, tolerance and iteration limit are those given here."""
    if cycle not in (0, 1, 2) or smoother != 0 or opchol != 0:
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "only cycle=0/1/2, smoother=0, opchol=0 are supported")
    if cycle != Ac.cycle:
        raise TmgError(_lib.TMG_ERR_ARG, "cycle differs from the hierarchy's (SetupGMG's mg.cycle)")
    if ngrids != Ac.ngrids:
        raise TmgError(_lib.TMG_ERR_ARG, "ngrids does not match the hierarchy")
    if (smoothiter1, smoothiter2, float(tolMG), maxiterMG) != Ac.params:
        raise TmgError(_lib.TMG_ERR_ARG, "SolveMG smoothing/tolerance/iteration arguments differ from SetupGMG's mg")
    xs, it, _, _ = Ac.solve(f, x0=x)
    return xs, it


def VCycle(A, b, x0, niter1, niter2, ngrids, nv, smoother, P, M, L, U, R, perm, opchol, opcond, pcgp=None):
    """x = VCycle(A,b,x0,...) (Implementation/VCycle.m:1-45) at the finest level, x0 = 0."""
    if nv != 1 or np.any(np.asarray(x0) != 0):
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "VCycle is exposed at nv = 1 with x0 = 0, as SolveMG calls it")
    if A.cycle != 0:
        raise TmgError(_lib.TMG_ERR_ARG, "the hierarchy was set up for W-cycles (use WCycle)")
    return A.vcycle(b)


def WCycle(A, b, x0, niter1, niter2, ngrids, nv, smoother, P, M, L, U, R, perm, opchol, opcond, pcgp=None):
    """x = WCycle(A,b,x0,...) (Implementation/WCycle.m:1-27) at the finest level, x0 = 0."""
    if nv != 1 or np.any(np.asarray(x0) != 0):
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "WCycle is exposed at nv = 1 with x0 = 0, as SolveMG calls it")
    if A.cycle != 1:
        raise TmgError(_lib.TMG_ERR_ARG, "the hierarchy was set up for V-cycles (use VCycle)")
    return A.vcycle(b)


def FMVCycle(A, b, niter1, niter2, ngrids, nv, smoother, P, M, L, U, R, perm, opchol, opcond, pcgp=None):
    """x = FMVCycle(A,b,niter1,niter2,ngrids,nv,...) (Implementation/FMVCycle.m:1-41) at the finest level."""
    if nv != 1:
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "FMVCycle is exposed at nv = 1, as SolveMG calls it")
    if A.cycle != 2:
        raise TmgError(_lib.TMG_ERR_ARG, "the hierarchy was not set up for FMV cycles (mg.cycle = 2)")
    return A.vcycle(b)


def Smooth(A, M, L, U, b, x0, niter, smoother, level=1):
    """x = Smooth(A{level},M{level},L,U,b,x0,niter,smoother) (Implementation/Smooth.m:1-24)."""
    if smoother != 0:
        raise TmgError(_lib.TMG_ERR_UNSUPPORTED, "only the Jacobi smoother is supported")
    return A.smooth(level, b, x0, niter)


def ElemNeighborhood(str, rmin, opfilter, fixedDens=None):
    """[W,w] = ElemNeighborhood(str,rmin,opfilter) (Implementation/ElemNeighborhood.m:1-100)."""
    fixed = None
    if fixedDens is not None and len(fixedDens):
        fixed = np.zeros(str.nelem, dtype=np.uint8)
        fixed[np.asarray(fixedDens, dtype=np.int64) - 1] = 1
    W = Filter(str.nelx, str.nely, str.nelz, rmin, opfilter, str.el, str.eh, str.ew, fixed)
    return W, W.rowsum()


def ApplyFilter(x, W, w, opfilter, trans):
    """xfil = ApplyFilter(x,W,w,opfilter,trans) (Implementation/ApplyFilter.m:1-21)."""
    if opfilter == 0:
        return np.asarray(x, dtype=np.float64).copy()
    return W.apply(x, trans=bool(trans))


def Grad(str, u, p, emin, k, x, Dofs, W, w, opfilter, mr=None, strDens=None, A=None, freeDensG=None):
    """Fgrad = Grad(str,u,p,emin,k,x,Dofs,W,w,opfilter,mr,strDens)
    (Implementation/Grad.m:1-64); with mr.op the gradient lives on the
    density mesh (Grad.m:24-49, mr.interp = false) and W is the density
    mesh's filter. ``A`` is the hierarchy whose handle runs the kernel (its
    p/emin and multiresolution factor must match)."""
    if _mr_factor(mr) != getattr(A, "mr", 1):
        raise TmgError(_lib.TMG_ERR_ARG, "mr does not match the hierarchy")
    if A is None:
        raise TmgError(_lib.TMG_ERR_ARG, "Grad needs the hierarchy handle A")
    freedens = None
    if freeDensG is not None:
        freedens = np.zeros(str.nelem, dtype=np.uint8)
        freedens[np.asarray(freeDensG, dtype=np.int64) - 1] = 1
    return A.grad(u, x, freedens=freedens, filt=W if opfilter != 0 else None)
