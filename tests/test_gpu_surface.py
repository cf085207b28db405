"""The reference's MATLAB call surface (tmg/gmg.py) driven in the order of
Implementation/StrSLP.m -- GlobalStiffMatrix, SetupGMG, SolveMG, Grad with
ElemNeighborhood / ApplyFilter, then a new design: GlobalStiffMatrix,
RefreshGMG and a warm-started SolveMG -- against the oracle, in single
resolution and with 2x2x2 density sub-elements; and the cycle functions
VCycle / WCycle / FMVCycle / Smooth against their oracle restatements."""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mg(L, cycle=0):
    return SimpleNamespace(ngrids=L, cycle=cycle, smoother=0, omega=0.6, smoothiter1=1, smoothiter2=1, tolMG=1e-8,
                           maxiterMG=5000, opchol=0, nmaxchol=60000)


ELEM = SimpleNamespace(deg=1, type=1)


@pytest.mark.parametrize("mrn", [1, 2])
def test_strslp_call_sequence(oracle_mod, mrn):
    import tmg
    _say(f"MATLAB surface sequence mr={mrn}")
    prob = P.cantilever(16, 8, 8)
    str_ = prob
    L = 3
    mr = SimpleNamespace(op=mrn == 2, n=2, d=2, interp=False)
    nd = prob.nelem * mrn ** 3
    dn = (16 * mrn, 8 * mrn, 8 * mrn)
    rng = np.random.default_rng(31)
    x1 = P.apply_filter(rng.uniform(0.05, 1.0, nd), *dn)
    x2 = P.apply_filter(np.clip(x1 + 0.05 * rng.standard_normal(nd), 0.01, 1.0), *dn)
    # device, through the MATLAB names
    K = tmg.GlobalStiffMatrix(str_, None, None, None, x1, 3.0, ELEM, 1e-9, mr if mrn == 2 else None)
    A, Pl, M, Lf, U, R, perm = tmg.SetupGMG(K, str_, _mg(L), ELEM)
    u1, it1 = tmg.SolveMG(np.zeros(prob.nfree), prob.f, A, Pl, M, Lf, U, R, perm, L, 0, 0, 1, 1, 1e-8, 5000, 0, 0)
    dstr = SimpleNamespace(nelx=dn[0], nely=dn[1], nelz=dn[2], nelem=nd, el=1.0 / mrn, eh=1.0 / mrn, ew=1.0 / mrn)
    W, w = tmg.ElemNeighborhood(dstr, 1.5 / mrn if mrn == 1 else 1.5, 2)
    g1 = tmg.Grad(str_, prob.full(u1), 3.0, 1e-9, None, x1, None, W, w, 2, mr if mrn == 2 else None, None, A=A)
    K2 = tmg.GlobalStiffMatrix(str_, None, None, None, x2, 3.0, ELEM, 1e-9, mr if mrn == 2 else None)
    tmg.RefreshGMG(A, K2)
    u2, it2 = tmg.SolveMG(u1, prob.f, A, Pl, M, Lf, U, R, perm, L, 0, 0, 1, 1, 1e-8, 5000, 0, 0)
    # oracle, the same sequence
    o = oracle_mod.Oracle(prob)
    (o.assemble(x1) if mrn == 1 else o.assemble_mr(2, x1))
    o.setup_gmg(L, 0.6)
    v1, jt1, _ = o.solve(prob.f)
    Wo, wo = oracle_mod.filter_matrix(*dn, 1.5 / mrn if mrn == 1 else 1.5, 2, el=1.0 / mrn, eh=1.0 / mrn,
                                      ew=1.0 / mrn)
    g0 = o.grad0(prob.full(u1), x1) if mrn == 1 else o.grad0_mr(2, prob.full(u1), x1)
    go = Wo.T @ (g0 / wo)
    (o.assemble(x2) if mrn == 1 else o.assemble_mr(2, x2))
    o.refresh()
    v2, jt2, _ = o.solve(prob.f, x0=v1)
    _say(f"  it {it1}/{jt1}, warm {it2}/{jt2}")
    assert abs(it1 - jt1) <= 1 and abs(it2 - jt2) <= 1
    assert _rel(u1, v1) < 1e-6 and _rel(u2, v2) < 1e-6
    assert np.abs(g1 - go).max() <= 1e-10 * np.abs(go).max()
    # ApplyFilter forward and transpose
    y = rng.standard_normal(nd)
    assert _rel(tmg.ApplyFilter(y, W, w, 2, False), (Wo @ y) / wo) < 1e-13
    assert _rel(tmg.ApplyFilter(y, W, w, 2, True), Wo.T @ (y / wo)) < 1e-13
    A.close()


@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_cycle_functions(oracle_mod, cycle):
    import tmg
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 33)
    L = 3
    K = tmg.GlobalStiffMatrix(prob, None, None, None, x, 3.0, ELEM, 1e-9)
    A, Pl, M, Lf, U, R, perm = tmg.SetupGMG(K, prob, _mg(L, cycle), ELEM)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(L, 0.6)
    b = np.random.default_rng(34).standard_normal(prob.nfree)
    z = np.zeros(prob.nfree)
    if cycle == 0:
        got, ref = tmg.VCycle(A, b, z, 1, 1, L, 1, 0, Pl, M, Lf, U, R, perm, 0, 0), o.vcycle(b)
    elif cycle == 1:
        got, ref = tmg.WCycle(A, b, z, 1, 1, L, 1, 0, Pl, M, Lf, U, R, perm, 0, 0), o.wcycle(b)
    else:
        got, ref = tmg.FMVCycle(A, b, 1, 1, L, 1, 0, Pl, M, Lf, U, R, perm, 0, 0), o.fmvcycle(b)
    assert _rel(got, ref) < 1e-8
    x0 = np.random.default_rng(35).standard_normal(prob.nfree)
    assert _rel(tmg.Smooth(A, M, Lf, U, b, x0, 2, 0), o.smooth(0, b, x0, 2)) < 1e-12
    # P{1} and its transpose
    xc = np.random.default_rng(36).standard_normal(o.nfree(1))
    assert _rel(Pl[0] @ xc, o.csr("P", 0) @ xc) < 1e-15
    xf = np.random.default_rng(37).standard_normal(prob.nfree)
    assert _rel(Pl[0].T @ xf, o.csr("P", 0).T @ xf) < 1e-14
    u, it = tmg.SolveMG(z, prob.f, A, Pl, M, Lf, U, R, perm, L, cycle, 0, 1, 1, 1e-8, 5000, 0, 0)
    v, jt, _ = o.solve(prob.f, cycle=cycle)
    assert abs(it - jt) <= 1 and _rel(u, v) < 1e-6
    A.close()
