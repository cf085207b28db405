"""Void-heavy designs (about half of the density elements exactly 0, so
E = Emin = 1e-9 there) taken from the oracle's own SLP loop at 48x24x24
(single resolution, 4 levels) and 24x12x12 with 2x2x2 density sub-elements
(3 levels), committed as tests/golden/void_heavy.{json,npz} by
make_golden.py --void. The device hierarchy must reproduce the oracle's
level operators on them, and the device solves its V-cycle and W-cycle
iteration counts and compliance: the regime in which the large
multiresolution SLP designs need many iterations is compared with the
reference's algorithm here."""
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


def _case(name):
    rec = json.load(open(os.path.join(HERE, "golden", "void_heavy.json")))[name]
    xfil = np.load(os.path.join(HERE, "golden", "void_heavy.npz"))[name + "_xfil"]
    return rec, P.cantilever(*rec["grid"]), xfil


@pytest.mark.parametrize("name", ["sr48", "mr24"])
def test_void_heavy_level_operators(name):
    import tmg
    rec, prob, xfil = _case(name)
    assert rec["void_fraction"] > 0.3
    s = tmg.Solver(prob, rec["levels"], precision="fp64", mr=rec["mr"])
    s.update_density(xfil)
    for lev, pr in enumerate(rec["level_probes"]):
        v = np.random.default_rng(100 + lev).standard_normal(s.nfree_level(lev + 1))
        Av = s.apply(lev + 1, v)
        _say(f"  {name} level {lev + 1}: ||Av|| {np.linalg.norm(Av):.12e} oracle {pr['norm_Av']:.12e}")
        assert abs(np.linalg.norm(Av) - pr["norm_Av"]) <= 1e-12 * pr["norm_Av"]
        assert abs(v @ Av - pr["vAv"]) <= 1e-11 * abs(pr["vAv"])
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("cycle", [0, 1])
@pytest.mark.parametrize("name", ["sr48", "mr24"])
def test_void_heavy_solves_match_oracle(name, cycle, precision):
    import tmg
    rec, prob, xfil = _case(name)
    ref = rec["V" if cycle == 0 else "W"]
    s = tmg.Solver(prob, rec["levels"], precision=precision, mr=rec["mr"], cycle=cycle)
    s.update_density(xfil)
    u, it, rr, rc = s.solve(prob.f)
    C = float(prob.f @ u)
    _say(f"  {name} cycle {cycle} {precision}: it {it} (oracle {ref['it']}) compliance err "
         f"{abs(C - ref['compliance']) / ref['compliance']:.2e}")
    assert rc == 0 and rr <= 1e-8
    assert abs(C - ref["compliance"]) <= 1e-7 * ref["compliance"]
    if precision == "fp64":
        assert abs(it - ref["it"]) <= 1
    else:
        assert it <= ref["it"] + 5
    s.close()
