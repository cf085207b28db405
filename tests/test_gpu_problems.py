"""Problems of the reference's own library (Implementation/Tests/*.mat),
rebuilt from the support, load and fixed-density rows recorded in
tests/golden/setup_fixtures.json (the GPU box has no reference tree), solved
on the device and checked against the FP64 oracle: a cantilever beam with a
point load, a two-load tower, a bridge with a load on an area and fixed
densities, and an L-shaped domain whose void region is fixed (gradient with
str.freeDensG and the filter decoupling the fixed elements)."""
import json
import os
import sys

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _fixture(name):
    for f in json.load(open(os.path.join(HERE, "golden", "setup_fixtures.json"))):
        if f["file"] == name:
            return f
    raise KeyError(name)


def _problem(name):
    fx = _fixture(name)
    prob = P.from_rows(fx["nelx"], fx["nely"], fx["nelz"], fx["supp"], fx["loads"], l=fx["l"], h=fx["h"], w=fx["w"],
                       fix=fx.get("fix"))
    assert prob.nfree == fx["nfree"] and prob.fixed_dens.size == fx["fixed_dens_count"]
    return prob


def _density(prob, seed):
    x = P.random_density(prob.nelx, prob.nely, prob.nelz, seed)
    if prob.fixed_dens.size:
        x[prob.fixed_dens - 1] = np.maximum(prob.fixed_dens_val, 1e-3)
    return x


CASES = [("cb48x16x16.mat", 4), ("tt16x80x16quarter.mat", 4), ("bd96x24x12quarter.mat", 2),
         ("ls48x48x16half.mat", 4)]


@pytest.mark.parametrize("name,L", CASES)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_problem_solves_match_oracle(oracle_mod, name, L, precision):
    import tmg
    _say(f"reference problem {name} {precision}")
    prob = _problem(name)
    x = _density(prob, 41)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(L, 0.6)
    u_or, it_or, _ = o.solve(prob.f)
    s = tmg.Solver(prob, L, precision=precision)
    s.update_density(x)
    u, it, rr, rc = s.solve(prob.f)
    r = prob.f - o.spmv(0, u)
    relres = np.linalg.norm(r) / np.linalg.norm(prob.f)
    C, Co = prob.f @ u, prob.f @ u_or
    _say(f"  it {it} (oracle {it_or}) true relres {relres:.2e} compliance err {abs(C - Co) / abs(Co):.2e}")
    assert rc == 0 and relres <= 1.2e-8 and abs(C - Co) / abs(Co) < 1e-7
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    s.close()


def test_fixed_region_gradient_matches_oracle(oracle_mod):
    """ls48x48x16half: 8,712 elements of the void region are fixed. The
    gradient is zero there (str.freeDensG = str.freeDens) and the filter
    decouples them (Implementation/ElemNeighborhood.m:90-96)."""
    import tmg
    _say("fixed-region gradient ls48x48x16half")
    prob = _problem("ls48x48x16half.mat")
    x = _density(prob, 42)
    mask = prob.free_dens_mask()
    fixed = 1 - mask
    s = tmg.Solver(prob, 4, precision="fp64")
    s.update_density(x)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0
    filt = tmg.Filter(prob.nelx, prob.nely, prob.nelz, 1.5, 2, fixed=fixed)
    g = s.grad(prob.full(u), x, freedens=mask, filt=filt)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    g0 = o.grad0(prob.full(u), x, freedens=mask)
    W, w = oracle_mod.filter_matrix(prob.nelx, prob.nely, prob.nelz, 1.5, 2, fixed=fixed)
    g_or = W.T @ (g0 / w)
    assert np.all(g0[mask == 0] == 0.0)
    assert np.abs(g - g_or).max() <= 1e-12 * np.abs(g_or).max()
    s.close()
