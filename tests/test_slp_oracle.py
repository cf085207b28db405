"""CPU pins of the SLP-loop restatement in oracle/slp_oracle.py
(Implementation/StrSLP.m:221-541) and of the product LP inside it.

The loop has no golden vectors in the reference (its LP is MATLAB's linprog);
it is pinned by what the algorithm fixes: the volume row is exact because the
filter is linear, accepted steps satisfy the merit test ared >= eta*pred,
rejected steps shrink the trust region as StrSLP.m:456-457 says, and the
product LP returns the same step as HiGHS on every instance the loop meets."""
import numpy as np
import pytest

from tmg import problem as P
from tmg.slp import lp_volume


@pytest.fixture(scope="module")
def slp_run(oracle_mod):
    import slp_oracle
    prob = P.cantilever(8, 4, 4)
    x0 = 0.3 + 0.05 * np.random.default_rng(5).standard_normal(prob.nelem)
    hist, x, xfil, u = slp_oracle.str_slp(prob, 0.3, ngrids=2, x0=x0, maxiter=8)
    return prob, hist, x, xfil, u


def test_slp_oracle_merit_and_trust_region(slp_run):
    prob, hist, x, xfil, u = slp_run
    assert len(hist) >= 8
    acc = [h for h in hist if h["accepted"]]
    assert len(acc) == 8
    for h in hist:
        if h["accepted"]:
            assert h["ared"] >= 0.1 * h["pred"]
        else:
            assert h["ared"] < 0.1 * h["pred"]
    # compliance of accepted points decreases from the start
    Fs = [h["F"] for h in acc]
    assert Fs[-1] < Fs[0]
    assert np.all((x >= -1e-12) & (x <= 1 + 1e-12))


def test_slp_oracle_volume_row_is_exact(slp_run):
    """b_new = b_old + A s: after the first step fixes feasibility the
    volume fraction stays at or below volfrac."""
    prob, hist, x, xfil, u = slp_run
    assert xfil.mean() <= 0.3 + 1e-10


def test_product_lp_equals_highs_on_slp_instances(oracle_mod):
    import slp_oracle
    prob = P.cantilever(8, 4, 4)
    o = oracle_mod.Oracle(prob)
    W, w = oracle_mod.filter_matrix(8, 4, 4, 1.5, 2)
    rng = np.random.default_rng(8)
    for k in range(4):
        x = np.clip(0.3 + 0.2 * rng.standard_normal(prob.nelem), 0, 1)
        xfil = (W @ x) / w
        o.assemble(xfil)
        if k == 0:
            o.setup_gmg(2)
        else:
            o.refresh()
        u, _, _ = o.solve(prob.f)
        g = W.T @ (o.grad0(prob.full(u), xfil) / w)
        velem = np.ones(prob.nelem)
        A = (W.T @ (velem / w)) / (0.3 * prob.nelem)
        b = (velem @ xfil - 0.3 * prob.nelem) / (0.3 * prob.nelem)
        delta = 0.1 * (k + 1)
        sL, sU = np.maximum(-delta, -x), np.minimum(delta, 1 - x)
        s_h, lam_h, ok_h = slp_oracle.lp_highs(g, A, -b, sL, sU)
        s, lam, ok = lp_volume(g, A, sL, sU, -b)
        assert ok == ok_h
        if ok:
            assert abs(g @ s - g @ s_h) <= 1e-10 * abs(g @ s_h)
            assert abs(lam - lam_h) <= 1e-8 * max(abs(lam_h), 1e-30) + 1e-300


@pytest.mark.parametrize("case", ["ineq_above", "eq_above", "eq_below", "eq_mixed_signs_below"])
def test_feasibility_step_matches_fallback_lp(oracle_mod, case):
    """The infeasible-LP fallback of StrSLP.m:271-285 (min t1 + t2 with
    [A -1 1] s = -b or <= -b) solved by HiGHS in the oracle, against the
    product's closed form: the least-violation box point, with the direction
    of the violation (an equality row whose target lies above every reachable
    volume must increase the volume, not decrease it)."""
    import slp_oracle
    from tmg.slp import feasibility_step
    rng = np.random.default_rng(["ineq_above", "eq_above", "eq_below", "eq_mixed_signs_below"].index(case) + 40)
    n = 400
    a = 0.5 + rng.random(n)
    if case == "eq_mixed_signs_below":
        a *= np.where(rng.random(n) < 0.3, -1.0, 1.0)
    x = rng.random(n)
    delta = 0.05
    sL2, sU2 = np.maximum(-0.8 * delta, -x), np.minimum(0.8 * delta, 1 - x)
    lo = float(np.where(a > 0, a * sL2, a * sU2).sum())
    hi = float(np.where(a > 0, a * sU2, a * sL2).sum())
    equality = case.startswith("eq")
    # b places the row target -b outside [lo, hi]
    b = -(lo - 5.0) if case.endswith("above") else -(hi + 5.0)
    s_ref, lam_ref = slp_oracle.fallback_lp_highs(a, b, sL2, sU2, equality=equality)
    s, lam = feasibility_step(a, b, sL2, sU2, equality=equality)
    np.testing.assert_allclose(s, s_ref, rtol=0, atol=1e-12)
    assert lam == pytest.approx(lam_ref, abs=1e-9)
    if case.endswith("below"):
        assert a @ s > 0 and lam == -1.0      # the volume grows toward the target
    else:
        assert a @ s < 0 and lam == 1.0


def test_feasibility_step_when_the_box_meets_the_row():
    from tmg.slp import feasibility_step
    rng = np.random.default_rng(4)
    a = 0.5 + rng.random(50)
    sL2, sU2 = -0.04 * np.ones(50), 0.04 * np.ones(50)
    for b, eq in ((0.2 * 0.04 * a.sum(), True), (-0.2 * 0.04 * a.sum(), True), (0.3, False), (-0.3, False)):
        s, lam = feasibility_step(a, b, sL2, sU2, equality=eq)
        assert lam == 0.0 and np.all(s >= sL2 - 1e-15) and np.all(s <= sU2 + 1e-15)
        if eq:
            assert abs(a @ s + b) <= 1e-12
        else:
            assert a @ s <= -b + 1e-12
