"""Sequential linear programming loop of the minimum-compliance problem on the
device solver (Implementation/StrSLP.m, GMG branch ``opsolver == 4``).

Every equilibrium solve, Galerkin refresh, density filter and compliance
gradient runs through libtmg on the GPU; the LP of each step is solved on the
host by ``tmg_lp_volume`` (a continuous knapsack: one volume row plus the
trust-region box, ``Implementation/StrSLP.m:221-279``). The trust-region,
merit-function and stopping logic follows ``Implementation/StrSLP.m:252-541``
step by step, so a run prints the same per-iteration line.

Not covered (they are not part of this solver's scope): element fixing and
degree elevation (``elem.fix``, ``Implementation/Elem2Fix.m``), the rounding
projection (``prj.op``), multiresolution with distinct density and design
meshes (``mr.n != mr.d``) and displacement interpolation (``mr.interp``).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import TMG_NOT_CONVERGED, TmgError, lib
from .gmg import Filter, Solver


@dataclass
class SLPParams:
    """The ``slp`` struct (Implementation/TopOpt3D.m:115-126)."""

    tolS: float = 1e-4
    tolG: float = 1e-3
    tolF: float = 5e-2
    maxcount: int = 3
    maxiter: int = 500
    delta0: float = 0.1
    eta: float = 0.1
    rho: float = 0.5
    alphaR: float = 0.25
    alphaA: float = 2.0


def lp_volume(c, a, sL, sU, r, equality=False):
    """``[s,fval,flag,~,lambda] = linprog(c, a', r, [], [], sL, sU)`` for one
    row (``Aeq`` form when ``equality``). Returns ``(s, lam, feasible)``."""
    c, a, sL, sU = (np.ascontiguousarray(v, dtype=np.float64) for v in (c, a, sL, sU))
    n = c.size
    if not (a.size == sL.size == sU.size == n):
        raise ValueError("LP vectors differ in length")
    s = np.empty(n)
    lam = C.c_double()
    dp = C.POINTER(C.c_double)
    rc = lib().tmg_lp_volume(n, c.ctypes.data_as(dp), a.ctypes.data_as(dp), sL.ctypes.data_as(dp),
                             sU.ctypes.data_as(dp), float(r), 1 if equality else 0, s.ctypes.data_as(dp),
                             C.byref(lam))
    if rc < 0:
        raise TmgError(rc, "tmg_lp_volume rejected its arguments")
    return s, lam.value, rc != TMG_NOT_CONVERGED


def feasibility_step(a, b, sL2, sU2, equality=False):
    """The step of the fallback LP the reference solves when the trust-region
    LP is infeasible (Implementation/StrSLP.m:271-285): minimise t1 + t2 over
    s in [sL2, sU2], t >= 0, subject to a's - t1 + t2 <= -b (= -b for the
    equality row). The optimum is the box point of least violation: the
    smallest a's when even that exceeds -b, and for the equality row the
    largest a's when even that falls short of -b. Returns (s, lam) with lam
    the row multiplier in linprog's sign convention (lambda.ineqlin >= 0;
    lambda.eqlin = +1 when the row is met from above, -1 from below).
    Coordinates with a = 0 take the zero step. When the box meets the row
    (no violation), the least-violation point is not unique; the step nearest
    to s = 0 along the row is returned with lam = 0."""
    a, sL2, sU2 = (np.asarray(v, dtype=np.float64) for v in (a, sL2, sU2))
    s_lo = np.where(a > 0, sL2, np.where(a < 0, sU2, 0.0))     # argmin a's
    s_hi = np.where(a > 0, sU2, np.where(a < 0, sL2, 0.0))     # argmax a's
    lo, hi = float(a @ s_lo), float(a @ s_hi)
    if lo > -b:
        return s_lo, 1.0
    if equality and hi < -b:
        return s_hi, -1.0
    # the row can be met inside the box: move from s = 0 toward the row
    # (scaled extreme step), which meets it exactly
    if (not equality and b <= 0) or b == 0:
        return np.zeros_like(a), 0.0
    t = -b / lo if b > 0 else -b / hi
    return t * (s_lo if b > 0 else s_hi), 0.0


@dataclass
class SLPResult:
    xstar: np.ndarray          # the last filtered densities computed (StrSLP.m:584)
    xstarDsgn: np.ndarray      # design variables
    u: np.ndarray              # full displacement vector
    opstop: int
    iter: int
    itrej: int
    F: float
    vol: float
    ns: int
    pcgit: list
    history: list = field(default_factory=list)
    time: dict = field(default_factory=dict)


def StrSLP(prob, volfrac, p=3.0, emin=1e-9, opfilter=2, rmin=1.5, volineq=True, slp=None, ngrids=4,
           precision="mixed", mr=1, x0=None, fixed_design=None, log=print, solver=None, filt=None, cycle=0):
    """Minimum compliance by SLP (Implementation/StrSLP.m:1-605) with the GMG
    solve on the device. ``mr = 2`` is multiresolution with ``mr.n = mr.d = 2``
    (densities and design variables on the same 2x finer mesh). ``x0`` is the
    initial design (default ``volfrac`` everywhere, Implementation/StartStrSLP.m:61).
    ``solver``/``filt`` may pass an existing hierarchy and filter to reuse;
    ``cycle`` is mg.cycle (0 V, 1 W, the reference's default, 2 FMV)."""
    slp = slp or SLPParams()
    tm = dict(setupK=0.0, setupPrec=0.0, system=0.0, grad=0.0, LP=0.0, filter=0.0)
    n = int(mr)
    dnx, dny, dnz = prob.nelx * n, prob.nely * n, prob.nelz * n
    nelem = dnx * dny * dnz
    L, H, Wd = prob.nelx * prob.el, prob.nely * prob.eh, prob.nelz * prob.ew
    V = L * H * Wd
    velem = np.full(nelem, V / nelem)
    Vmax = volfrac * V
    fixed_val = np.zeros(nelem)
    if fixed_design is None and n == 1 and getattr(prob, "fixed_dens", None) is not None and prob.fixed_dens.size:
        # the problem's void / solid regions (str.fixedDens, str.fixedDensVal)
        fixed = np.zeros(nelem, dtype=bool)
        fixed[prob.fixed_dens - 1] = True
        fixed_val[prob.fixed_dens - 1] = prob.fixed_dens_val
    else:
        fixed = np.zeros(nelem, dtype=bool) if fixed_design is None else np.asarray(fixed_design, dtype=bool)
    freeD = ~fixed
    anyfixed = bool(fixed.any())

    def free(v):
        """v on the free design variables (v itself when none is fixed)."""
        return v[freeD] if anyfixed else v

    def box_gap(xv, g):
        """max |min(1, max(0, x - g)) - x| over the free design variables
        (StrSLP.m:247-256, 482-486), with one temporary."""
        xf = free(xv)
        if xf.size == 0:
            return 0.0
        t = xf - g
        np.maximum(t, 0.0, out=t)
        np.minimum(t, 1.0, out=t)
        t -= xf
        np.abs(t, out=t)
        return float(t.max())

    def trust_box(d, scale=1.0):
        """Box bounds of the step: max(-d, -x) <= s <= min(d, 1 - x), 0 on
        fixed design variables (StrSLP.m:234-238)."""
        lo = np.negative(x)
        np.maximum(lo, -scale * d, out=lo)
        hi = np.subtract(1.0, x)
        np.minimum(hi, scale * d, out=hi)
        if anyfixed:
            lo[fixed] = 0.0
            hi[fixed] = 0.0
        return lo, hi

    # str.freeDensG = str.freeDens: fixed elements carry no gradient (StrSLP.m:75, Grad.m:18-23)
    freedens = freeD.astype(np.uint8) if (fixed.any() and n == 1) else None

    t0 = time.perf_counter()
    own_filter = filt is None
    if opfilter != 0 and own_filter:
        filt = Filter(dnx, dny, dnz, rmin, opfilter, el=L / dnx, eh=H / dny, ew=Wd / dnz,
                      fixed=fixed.astype(np.uint8) if fixed.any() else None)
    own_solver = solver is None
    if own_solver:
        solver = Solver(prob, ngrids, precision=precision, p=p, emin=emin, mr=n, cycle=cycle)
    tm["setupPrec"] += time.perf_counter() - t0

    def apply_filter(v, trans=False):
        t = time.perf_counter()
        out = filt.apply(v, trans) if opfilter != 0 else np.array(v, dtype=np.float64)
        tm["filter"] += time.perf_counter() - t
        return out

    def analyse(xfil, uold):
        t = time.perf_counter()
        solver.update_density(xfil)          # GlobalStiffMatrix + Galerkin refresh (StrSLP.m:319,345-365)
        tm["setupPrec"] += time.perf_counter() - t
        t = time.perf_counter()
        u_free, it, _, _ = solver.solve(prob.f, uold)
        tm["system"] += time.perf_counter() - t
        return u_free, it

    def gradient(u_full, xfil):
        t = time.perf_counter()
        g = solver.grad(u_full, xfil, freedens=freedens, filt=filt if opfilter != 0 else None)
        tm["grad"] += time.perf_counter() - t
        return g

    if x0 is None:
        # Implementation/StartStrSLP.m:61-64
        x = np.full(nelem, float(volfrac))
        if fixed.any():
            x[fixed] = fixed_val[fixed]
            x[freeD] = (volfrac * V - fixed_val[fixed].sum() * (L / dnx) * (H / dny) * (Wd / dnz)) / freeD.sum()
    else:
        x = np.array(x0, dtype=np.float64)
    delta = slp.delta0
    theta1 = theta2 = thetaMax = 1.0
    countS = countG = countF = 0
    iter_ = itrej = 0
    pcgit = []
    history = []

    xfil = apply_filter(x)
    u_free, it = analyse(xfil, np.zeros(prob.nfree))
    pcgit.append(it)
    ns = 1
    Fold = float(prob.f @ u_free)
    Fgrad = gradient(prob.full(u_free), xfil)

    # volume row and its right-hand side (StrSLP.m:226-233)
    A = (apply_filter(velem, True) if opfilter != 0 else velem) / Vmax
    b = (velem @ xfil - Vmax) / Vmax
    sL, sU = trust_box(delta)

    # optimality of the initial point (StrSLP.m:247-256)
    lam = 0.0
    F = Fold
    gpnorm = box_gap(x, free(Fgrad))
    if gpnorm < 1e-12 and b < 1e-12 and (x - 1.0).max() < 1e-12 and x.min() > -1e-12:
        countG = countF = slp.maxcount

    while (countG < slp.maxcount or countF < slp.maxcount) and countS < slp.maxcount and iter_ < slp.maxiter:
        itno = iter_ + 1
        t = time.perf_counter()
        s, lam, ok = lp_volume(Fgrad, A, sL, sU, -b, equality=not volineq)
        flag = 1
        if not ok:
            # infeasible trust-region LP: minimise the volume violation instead (StrSLP.m:271-285)
            flag = -2
            sL2, sU2 = trust_box(delta, 0.8)
            s, lam = feasibility_step(A, b, sL2, sU2, equality=not volineq)
        fval = float(Fgrad @ s)
        tm["LP"] += time.perf_counter() - t

        snorm = float(np.abs(s).max()) if s.size else 0.0
        xold, bold, deltaold = x, b, delta
        x = x + s
        xfil = apply_filter(x)
        b = (velem @ xfil - Vmax) / Vmax

        uold = u_free.copy()            # pcgp.opu0 = true: warm start
        u_new, it = analyse(xfil, uold)
        pcgit.append(it)
        ns += 1
        F = float(prob.f @ u_new)

        predopt = -fval
        predfsb = abs(bold) - abs(float(A @ s) + bold)
        aredopt = Fold - F
        aredfsb = abs(bold) - abs(b)
        thetaMin = min(theta1, theta2)
        thetaLarge = (1 + 1e6 / ((iter_ + 1) ** 1.1)) * thetaMin
        thetaSup = 1.0 if predopt > 0.5 * predfsb else (0.5 * predfsb) / (predfsb - predopt)
        theta = min(thetaLarge, thetaSup, thetaMax)
        theta2, theta1 = theta1, theta
        pred = theta * predopt + (1 - theta) * predfsb
        ared = theta * aredopt + (1 - theta) * aredfsb

        accepted = ared >= slp.eta * pred
        if not accepted:
            itrej += 1
            delta = max(slp.alphaR * snorm, 0.1 * deltaold)
            thetaMax = theta
            x, b, F = xold, bold, Fold
            # the rejected solve's displacement is not kept as the next warm
            # start: StrSLP.m:329-333 reads u, which the rejected solve overwrote
            u_free = u_new
        else:
            if ared >= slp.rho * pred:
                delta = min(slp.alphaA * deltaold, 1.0)
            delta = max(delta, 0.0001)
            thetaMax = 1.0
            Fold = F
            u_free = u_new
            Fgrad = gradient(prob.full(u_free), xfil)
            Lgrad = free(A) * lam
            Lgrad += free(Fgrad)
            gpnorm = box_gap(x, Lgrad)
            iter_ += 1
            countG = countG + 1 if gpnorm <= slp.tolG else 0
            countF = countF + 1 if ared <= slp.tolF else 0
        countS = countS + 1 if snorm <= slp.tolS else 0

        sL, sU = trust_box(delta)
        rec = dict(it=itno, F=F, ared=ared, pred=pred, snorm=snorm, delta=deltaold, gpnorm=gpnorm, flag=flag,
                   pcgit=it, accepted=bool(accepted), lam=lam)
        history.append(rec)
        if log:
            log(f"It: {itno:3d} | F: {F:+e} | Ared: {ared:+e} | Pred: {pred:+e} | snorm: {snorm:f} | "
                f"delta: {deltaold:f} | gpnorm: {gpnorm:f} | flagLP: {flag:d} | pcgit: {it:4d}")

    if countG >= slp.maxcount and countF >= slp.maxcount:
        opstop = 0
    elif countS >= slp.maxcount:
        opstop = 1
    else:
        opstop = 2
    # as the reference returns it (StrSLP.m:584): the last filtered densities
    # computed, which belong to the rejected trial point if the last step was rejected
    xstar = xfil
    vol = float(velem @ xstar) / V
    res = SLPResult(xstar=xstar, xstarDsgn=x, u=prob.full(u_free), opstop=opstop, iter=iter_, itrej=itrej, F=F,
                    vol=vol, ns=ns, pcgit=pcgit, history=history, time=tm)
    if own_solver:
        solver.close()
    if own_filter and filt is not None:
        filt.close()
    return res
