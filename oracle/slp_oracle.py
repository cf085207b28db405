"""CPU restatement of the SLP loop (test infrastructure only: the product
never imports this module).

It follows Implementation/StrSLP.m for the assembled GMG branch
(opsolver = 4, genK = true) with no element fixing and no rounding
projection, on the FP64 oracle:

* analysis: GlobalStiffMatrix + SetupGMG / Galerkin refresh + SolveMG
  (StrSLP.m:79-110, 316-369) through ``oracle.Oracle``;
* objective F = f'u (StrSLP.m:166, 424) and gradient
  Fgrad = W'(g0./w) (Implementation/Grad.m:18-62, ApplyFilter.m:15-16);
* the LP (StrSLP.m:221-285) by scipy's HiGHS dual simplex, the same method
  family as the reference's ``linprog(...,'dual-simplex')``;
* merit function, trust region and stopping tests (StrSLP.m:427-541).

Parity of the LP itself is pinned separately (KKT conditions and HiGHS on
random instances); this module pins the loop: the sequence of objective
values, step acceptances and trust-region radii.
"""
from __future__ import annotations

import numpy as np

import oracle


def lp_highs(c, A, rhs, sL, sU, equality=False):
    """[s,fval,flag,~,lambda] = linprog(c,A,rhs,[],[],sL,sU) for one row A (StrSLP.m:265-271)."""
    from scipy.optimize import linprog
    bounds = np.column_stack([sL, sU])
    if equality:
        r = linprog(c, A_eq=A[None, :], b_eq=[rhs], bounds=bounds, method="highs-ds")
        lam = -r.eqlin.marginals[0] if r.status == 0 else 0.0
    else:
        r = linprog(c, A_ub=A[None, :], b_ub=[rhs], bounds=bounds, method="highs-ds")
        lam = -r.ineqlin.marginals[0] if r.status == 0 else 0.0
    return (r.x if r.status == 0 else None), float(lam), r.status == 0


def fallback_lp_highs(A, b, sL2, sU2, equality=False):
    """The reference's LP for an infeasible trust-region step
    (StrSLP.m:271-285): variables [s; t1; t2], cost [0; 1; 1], row
    [A -1 1] (<= or = -b), bounds [sL2; 0; 0] .. [sU2; inf; inf], solved by
    HiGHS' dual simplex as stated. Returns (s, lambda of the row)."""
    n = A.size
    c2 = np.concatenate([np.zeros(n), [1.0, 1.0]])
    A2 = np.concatenate([A, [-1.0, 1.0]])
    lo = np.concatenate([sL2, [0.0, 0.0]])
    hi = np.concatenate([sU2, [np.inf, np.inf]])
    s2, lam, ok = lp_highs(c2, A2, -b, lo, hi, equality=equality)
    assert ok, "the fallback LP is always feasible"
    return s2[:n], lam


def str_slp(prob, volfrac, p=3.0, emin=1e-9, rmin=1.5, opfilter=2, ngrids=3, mr=1, x0=None, maxiter=10,
            delta0=0.1, eta=0.1, rho=0.5, alphaR=0.25, alphaA=2.0, tolS=1e-4, tolG=1e-3, tolF=5e-2,
            maxcount=3):
    """Returns (history, x, xfil, u_full) with one history record per LP solved."""
    n = int(mr)
    dnx, dny, dnz = prob.nelx * n, prob.nely * n, prob.nelz * n
    nelem = dnx * dny * dnz
    Ltot, Htot, Wtot = prob.nelx * prob.el, prob.nely * prob.eh, prob.nelz * prob.ew
    V = Ltot * Htot * Wtot
    velem = np.full(nelem, V / nelem)                       # StrSLP.m:52-58
    Vmax = volfrac * V
    W, w = oracle.filter_matrix(dnx, dny, dnz, rmin, opfilter, el=Ltot / dnx, eh=Htot / dny, ew=Wtot / dnz)
    filt = lambda v: (W @ v) / w                            # ApplyFilter.m:18
    filtT = lambda v: W.T @ (v / w)                         # ApplyFilter.m:15-16

    o = oracle.Oracle(prob, p=p, Emin=emin)

    def analyse(xfil, first, x0v):
        if n == 1:
            o.assemble(xfil)
        else:
            o.assemble_mr(n, xfil)
        if first:
            o.setup_gmg(ngrids, 0.6)                        # SetupGMG (StrSLP.m:98)
        else:
            o.refresh()                                     # Galerkin refresh (StrSLP.m:345-365)
        u, it, _ = o.solve(prob.f, x0v)
        return u, it

    def gradient(u, xfil):
        uf = prob.full(u)
        g0 = o.grad0(uf, xfil) if n == 1 else o.grad0_mr(n, uf, xfil)
        return filtT(g0)

    x = np.full(nelem, float(volfrac)) if x0 is None else np.array(x0, dtype=np.float64)
    delta = delta0
    theta1 = theta2 = thetaMax = 1.0
    countS = countG = countF = 0
    it_acc = itrej = 0
    xfil = filt(x)
    u, it = analyse(xfil, True, np.zeros(prob.nfree))
    Fold = float(prob.f @ u)
    Fgrad = gradient(u, xfil)
    A = filtT(velem) / Vmax                                  # StrSLP.m:226-231
    b = (velem @ xfil - Vmax) / Vmax                        # StrSLP.m:233
    sL, sU = np.maximum(-delta, -x), np.minimum(delta, 1 - x)
    hist = []
    F = Fold
    gpnorm = 0.0
    while (countG < maxcount or countF < maxcount) and countS < maxcount and it_acc < maxiter:
        s, lam, ok = lp_highs(Fgrad, A, -b, sL, sU)
        if not ok:                                          # StrSLP.m:271-285
            sL2, sU2 = np.maximum(-0.8 * delta, -x), np.minimum(0.8 * delta, 1 - x)
            s, lam = fallback_lp_highs(A, b, sL2, sU2)
        fval = float(Fgrad @ s)
        snorm = float(np.abs(s).max())
        xold, bold, deltaold = x, b, delta
        x = x + s
        xfil = filt(x)
        b = (velem @ xfil - Vmax) / Vmax
        u, it = analyse(xfil, False, u)                     # warm start, pcgp.opu0 (StrSLP.m:329-333)
        F = float(prob.f @ u)
        predopt = -fval
        predfsb = abs(bold) - abs(float(A @ s) + bold)
        aredopt = Fold - F
        aredfsb = abs(bold) - abs(b)
        thetaMin = min(theta1, theta2)
        thetaLarge = (1 + 1e6 / ((it_acc + 1) ** 1.1)) * thetaMin
        thetaSup = 1.0 if predopt > 0.5 * predfsb else (0.5 * predfsb) / (predfsb - predopt)
        theta = min(thetaLarge, thetaSup, thetaMax)
        theta2, theta1 = theta1, theta
        pred = theta * predopt + (1 - theta) * predfsb
        ared = theta * aredopt + (1 - theta) * aredfsb
        accepted = ared >= eta * pred
        if not accepted:                                    # StrSLP.m:455-462
            itrej += 1
            delta = max(alphaR * snorm, 0.1 * deltaold)
            thetaMax = theta
            x, b, F = xold, bold, Fold
        else:                                               # StrSLP.m:463-485
            if ared >= rho * pred:
                delta = min(alphaA * deltaold, 1.0)
            delta = max(delta, 0.0001)
            thetaMax = 1.0
            Fold = F
            Fgrad = gradient(u, xfil)
            Lgrad = Fgrad + lam * A
            gpnorm = float(np.abs(np.minimum(1, np.maximum(0.0, x - Lgrad)) - x).max())
            it_acc += 1
            countG = countG + 1 if gpnorm <= tolG else 0
            countF = countF + 1 if ared <= tolF else 0
        countS = countS + 1 if snorm <= tolS else 0
        sL, sU = np.maximum(-delta, -x), np.minimum(delta, 1 - x)
        hist.append(dict(F=F, ared=ared, pred=pred, snorm=snorm, delta=deltaold, gpnorm=gpnorm, accepted=accepted,
                         pcgit=it, lam=lam))
    return hist, x, xfil, prob.full(u)
