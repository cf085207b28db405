"""The 192x96x96 cantilever (5.4M free DOFs, filtered random densities seed 3,
5-level GMG V(1,1), tol 1e-8) in FP64 and mixed precision against the FP64
oracle's solve of the same problem, committed as a fixture
(tests/golden/oracle_192x96x96_seed3.json, minted by make_golden.py --only
192x96x96 in the build container: compliance, iteration count, ||u|| and a
fixed sample of 353 entries of u)."""
import json
import os
import sys

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def config3():
    g = json.load(open(os.path.join(HERE, "golden", "oracle_192x96x96_seed3.json")))
    prob = P.cantilever(192, 96, 96)
    x = P.random_density(192, 96, 96, 3)
    assert prob.nfree == g["nfree"]
    return g, prob, x


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_against_oracle_fixture(config3, precision):
    import tmg
    g, prob, x = config3
    _say(f"config3 {precision}")
    s = tmg.Solver(prob, 5, precision=precision)
    s.update_density(x)
    u, it, rr, rc = s.solve(prob.f)
    C = float(prob.f @ u)
    idx = np.asarray(g["u_sample_index"])
    us = np.asarray(g["u_sample"])
    err_C = abs(C - g["compliance"]) / abs(g["compliance"])
    err_u = np.linalg.norm(u[idx] - us) / np.linalg.norm(us)
    err_norm = abs(np.linalg.norm(u) - g["u_norm"]) / g["u_norm"]
    # true residual with the FP64 operator of an FP64 handle (matrix-free
    # K(rho) u, itself equal to the oracle's assembled operator at every
    # level in test_gpu_parity)
    s64 = s if precision == "fp64" else tmg.Solver(prob, 5, precision="fp64")
    if s64 is not s:
        s64.update_density(x)
    relres = np.linalg.norm(prob.f - s64.apply(1, u)) / np.linalg.norm(prob.f)
    _say(f"  it {it} (oracle {g['it']}) relres {rr:.3e} true {relres:.3e} compliance err {err_C:.2e} "
         f"u sample err {err_u:.2e} ||u|| err {err_norm:.2e}")
    assert rc == 0 and rr <= 1e-8
    assert relres <= 1.2e-8
    assert err_C < 1e-7
    assert err_u < 1e-6 and err_norm < 1e-6
    if precision == "fp64":
        assert abs(it - g["it"]) <= 1
    else:
        assert it <= g["it"] + 3
    s.close()
