"""Device SLP loop (tmg.slp.StrSLP: solve, Galerkin refresh, filter and
gradient on the GPU, LP on the host) against the oracle's restatement of
Implementation/StrSLP.m on the same inputs: the same sequence of accepted and
rejected steps, objective values within the solve tolerance's reach, and the
same final design."""
import sys

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision,mr", [("fp64", 1), ("mixed", 1), ("fp64", 2), ("mixed", 2)])
def test_slp_loop_matches_oracle(oracle_mod, precision, mr):
    _say(f"SLP loop {precision} mr={mr}")
    import slp_oracle
    from tmg import slp
    prob = P.cantilever(8, 4, 4)
    n = prob.nelem * mr ** 3
    x0 = 0.3 + 0.05 * np.random.default_rng(5).standard_normal(n)
    iters = 8 if mr == 1 else 5
    hist_o, x_o, xfil_o, u_o = slp_oracle.str_slp(prob, 0.3, ngrids=2, mr=mr, x0=x0, maxiter=iters)
    res = slp.StrSLP(prob, 0.3, ngrids=2, precision=precision, mr=mr, x0=x0,
                     slp=slp.SLPParams(maxiter=iters), log=None)
    hist = res.history
    assert [h["accepted"] for h in hist] == [bool(h["accepted"]) for h in hist_o]
    tolF = 1e-7 if precision == "fp64" else 1e-6
    for k, (h, ho) in enumerate(zip(hist, hist_o)):
        _say(f"  step {k}: F {h['F']:.9e} oracle {ho['F']:.9e} pcgit {h['pcgit']} / {ho['pcgit']}")
        assert abs(h["F"] - ho["F"]) <= tolF * abs(ho["F"])
        assert abs(h["delta"] - ho["delta"]) <= 1e-12
        if precision == "fp64":
            assert abs(h["pcgit"] - ho["pcgit"]) <= 1
    assert np.abs(res.xstarDsgn - x_o).max() <= 1e-5
    assert np.abs(res.xstar - xfil_o).max() <= 1e-5
    assert res.iter == iters
    assert res.vol <= 0.3 + 1e-9
