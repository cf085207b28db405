"""The SLP step's LP (tmg_lp_volume, replacing linprog at
Implementation/StrSLP.m:265-285): checked against scipy's HiGHS dual simplex and
against the KKT conditions of the one-row box LP. Host code only: no GPU."""
import time

import numpy as np
import pytest

from tmg.slp import lp_volume


def _instance(n, seed, tie=False):
    rng = np.random.default_rng(seed)
    c = -rng.random(n) if not tie else -np.repeat(rng.random(max(n // 8, 1)), 8)[:n]
    a = 0.5 + rng.random(n)
    if tie:
        a = np.ones(n)
    x = rng.random(n)
    delta = 0.2
    sL, sU = np.maximum(-delta, -x), np.minimum(delta, 1 - x)
    return c, a, sL, sU, x


def _highs(c, a, sL, sU, r, equality):
    from scipy.optimize import linprog
    kw = dict(A_eq=a[None, :], b_eq=[r]) if equality else dict(A_ub=a[None, :], b_ub=[r])
    res = linprog(c, bounds=np.column_stack([sL, sU]), method="highs-ds", **kw)
    assert res.status == 0
    m = res.eqlin.marginals[0] if equality else res.ineqlin.marginals[0]
    return res.x, float(res.fun), -float(m)


def _kkt(c, a, sL, sU, r, s, lam, equality, tol=1e-10):
    assert np.all(s >= sL - tol) and np.all(s <= sU + tol)
    g = a @ s
    if equality:
        assert abs(g - r) <= tol * max(1.0, abs(r))
    else:
        assert g <= r + tol * max(1.0, abs(r))
        assert lam >= 0
        assert lam * (r - g) <= 1e-9 * max(1.0, abs(r))           # complementary slackness
    d = c + lam * a                                              # reduced costs
    inside = (s > sL + 1e-12) & (s < sU - 1e-12)
    assert np.all(np.abs(d[inside]) <= 1e-9 * (1 + np.abs(c[inside])))
    assert np.all(d[(s <= sL + 1e-12) & (sL < sU)] >= -1e-9)
    assert np.all(d[(s >= sU - 1e-12) & (sL < sU)] <= 1e-9)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("equality", [False, True])
def test_lp_matches_highs(seed, equality):
    n = 50 + 37 * seed
    c, a, sL, sU, _ = _instance(n, seed)
    # a row that binds: between the smallest and largest a's over the box
    lo, hi = a @ sL, a @ sU
    r = lo + (0.2 + 0.1 * seed) * (hi - lo)
    s, lam, ok = lp_volume(c, a, sL, sU, r, equality)
    assert ok
    xs, fun, lam_h = _highs(c, a, sL, sU, r, equality)
    assert abs(c @ s - fun) <= 1e-11 * max(1.0, abs(fun))
    assert abs(lam - lam_h) <= 1e-9 * max(1.0, abs(lam_h))
    assert np.abs(s - xs).max() <= 1e-9                          # distinct breakpoints: unique optimum
    _kkt(c, a, sL, sU, r, s, lam, equality)


def test_lp_slack_row_gives_zero_multiplier():
    c, a, sL, sU, _ = _instance(80, 11)
    r = a @ sU + 1.0                                             # never binds
    s, lam, ok = lp_volume(c, a, sL, sU, r)
    assert ok and lam == 0.0
    assert np.array_equal(s, np.where(c < 0, sU, sL))


def test_lp_tied_breakpoints():
    c, a, sL, sU, _ = _instance(96, 5, tie=True)
    r = a @ sL + 0.37 * (a @ sU - a @ sL)
    s, lam, ok = lp_volume(c, a, sL, sU, r)
    _, fun, lam_h = _highs(c, a, sL, sU, r, False)
    assert ok and abs(c @ s - fun) <= 1e-11 * abs(fun) and abs(lam - lam_h) <= 1e-9
    _kkt(c, a, sL, sU, r, s, lam, False)


def test_lp_infeasible_returns_least_violating_point():
    c, a, sL, sU, _ = _instance(64, 2)
    r = a @ sL - 1.0
    s, lam, ok = lp_volume(c, a, sL, sU, r)
    assert not ok
    assert np.array_equal(s, sL)
    s, lam, ok = lp_volume(c, a, sL, sU, a @ sU + 1.0, equality=True)
    assert not ok and np.array_equal(s, sU)


def test_lp_mixed_sign_rows_and_zero_coefficients():
    rng = np.random.default_rng(9)
    n = 120
    c = rng.standard_normal(n)
    a = rng.standard_normal(n)
    a[::10] = 0.0
    sL, sU = -rng.random(n), rng.random(n)
    for eq in (False, True):
        r = 0.3 * np.abs(a) @ (sU - sL) * (0.1 if eq else -0.05)
        s, lam, ok = lp_volume(c, a, sL, sU, r, eq)
        xs, fun, lam_h = _highs(c, a, sL, sU, r, eq)
        assert ok and abs(c @ s - fun) <= 1e-10 * max(1.0, abs(fun))
        assert abs(lam - lam_h) <= 1e-8 * max(1.0, abs(lam_h))
        _kkt(c, a, sL, sU, r, s, lam, eq)


def test_lp_edge_cases_and_errors():
    s, lam, ok = lp_volume(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), 0.0)
    assert ok and s.size == 0 and lam == 0.0
    with pytest.raises(Exception):
        lp_volume(np.ones(3), np.ones(3), np.ones(3), np.zeros(3), 0.0)      # sL > sU
    with pytest.raises(ValueError):
        lp_volume(np.ones(3), np.ones(2), np.zeros(3), np.ones(3), 0.0)


def test_lp_large_instance_is_linear_time():
    n = 3_000_000
    c, a, sL, sU, _ = _instance(n, 4)
    r = a @ sL + 0.3 * (a @ sU - a @ sL)
    t = time.perf_counter()
    s, lam, ok = lp_volume(c, a, sL, sU, r)
    dt = time.perf_counter() - t
    assert ok
    _kkt(c, a, sL, sU, r, s, lam, False, tol=1e-9)
    assert dt < 5.0


def test_lp_thread_count_does_not_change_the_step(monkeypatch):
    n = 1_500_000
    c, a, sL, sU, _ = _instance(n, 12)
    r = a @ sL + 0.45 * (a @ sU - a @ sL)
    out = []
    for nt in ("1", "3", "16"):
        monkeypatch.setenv("TMG_LP_THREADS", nt)
        out.append(lp_volume(c, a, sL, sU, r))
    for s, lam, ok in out[1:]:
        assert ok and lam == out[0][1]
        # every bound is identical; the one fractional element absorbs the
        # rounding of the row sum (1.5M terms summed in per-thread order)
        diff = np.flatnonzero(s != out[0][0])
        assert diff.size <= 1
        assert np.abs(s - out[0][0]).max() <= 1e-7


def test_lp_large_tie_groups_across_threads(monkeypatch):
    """8 distinct breakpoints shared by 200k elements each: the fractional
    weight is spread over a tie group that spans several threads' chunks."""
    monkeypatch.setenv("TMG_LP_THREADS", "8")
    n = 1_600_000
    rng = np.random.default_rng(13)
    c = -np.repeat(np.linspace(1.0, 2.0, 8), n // 8)
    a = np.ones(n)
    sL, sU = -0.1 * rng.random(n), 0.1 * rng.random(n)
    r = a @ sL + 0.37 * (a @ sU - a @ sL)
    s, lam, ok = lp_volume(c, a, sL, sU, r)
    assert ok
    _kkt(c, a, sL, sU, r, s, lam, False, tol=1e-9)
    frac = (s > sL + 1e-12) & (s < sU - 1e-12)
    assert np.all(-c[frac] == lam)
    # optimal value by the greedy closed form: fill the cheapest breakpoints first
    order = np.argsort(c, kind="stable")
    budget = r - a @ sL
    take = np.minimum(np.maximum(budget - np.concatenate([[0.0], np.cumsum((sU - sL)[order])[:-1]]), 0),
                      (sU - sL)[order])
    best = c @ sL + c[order] @ take
    assert abs(c @ s - best) <= 1e-9 * abs(best)
