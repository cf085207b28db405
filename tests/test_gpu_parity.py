"""GPU parity of libtmg against the FP64 oracle (the reference's assembled
GMG-PCG path restated on the CPU), through the C-ABI."""
import sys

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(prob, x, L, precision="fp64", **kw):
    import tmg
    s = tmg.Solver(prob, L, precision=precision, **kw)
    s.update_density(x)
    return s


def _oracle(oracle_mod, prob, x, L):
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(L, 0.6)
    return o


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


SMALL = [(8, 4, 4, 3, 11), (16, 8, 8, 3, 12), (12, 8, 4, 2, 13)]


@pytest.mark.parametrize("nelx,nely,nelz,L,seed", SMALL)
def test_level_operators_match_oracle(oracle_mod, nelx, nely, nelz, L, seed):
    _say(f"level operators {nelx}x{nely}x{nelz} L={L}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    s = _setup(prob, x, L)
    rng = np.random.default_rng(seed)
    for lev in range(1, L + 1):
        nel, fd = s.level_info(lev)
        assert nel == o.level_dims(lev - 1)
        np.testing.assert_array_equal(fd - 1, o.level_free(lev - 1))
        v = rng.standard_normal(fd.size)
        yo = o.spmv(lev - 1, v)
        assert _rel(s.apply(lev, v), yo) < 1e-13, lev
        # smoother (Implementation/Smooth.m) from a random x0, 2 sweeps
        b = rng.standard_normal(fd.size)
        assert _rel(s.smooth(lev, b, v, 2), o.smooth(lev - 1, b, v, 2)) < 1e-12
        if lev < L:
            Pm = o.csr("P", lev - 1)
            vc = rng.standard_normal(Pm.shape[1])
            assert _rel(s.prolong(lev, vc), Pm @ vc) < 1e-15
            vf = rng.standard_normal(Pm.shape[0])
            assert _rel(s.restrict(lev, vf), Pm.T @ vf) < 1e-14


@pytest.mark.parametrize("nelx,nely,nelz,L,seed", SMALL)
def test_vcycle_matches_oracle(oracle_mod, nelx, nely, nelz, L, seed):
    _say(f"vcycle {nelx}x{nely}x{nelz}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    s = _setup(prob, x, L)
    b = np.random.default_rng(seed).standard_normal(prob.nfree)
    # the coarse solve is an explicit FP64 inverse vs the oracle's triangular
    # solves: agreement to the coarse system's conditioning times eps
    assert _rel(s.vcycle(b), o.vcycle(b)) < 1e-8


def _check_solution(oracle_mod, o, prob, u, it_gpu, it_or, u_or, tol_u):
    r = prob.f - o.spmv(0, u)
    relres = np.linalg.norm(r) / np.linalg.norm(prob.f)
    C, Co = prob.f @ u, prob.f @ u_or
    _say(f"  it {it_gpu} (oracle {it_or}) true relres {relres:.3e} compliance rel err {abs(C - Co) / abs(Co):.3e}")
    assert relres <= 1.2e-8
    assert abs(C - Co) / abs(Co) < 1e-7
    assert _rel(u, u_or) < tol_u
    return relres


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz,L,seed", SMALL)
def test_solve_small(oracle_mod, precision, nelx, nely, nelz, L, seed):
    _say(f"solve {precision} {nelx}x{nely}x{nelz}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    u_or, it_or, _ = o.solve(prob.f)
    s = _setup(prob, x, L, precision)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    else:
        assert it <= it_or + 3
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz,L,seed", [(16, 8, 8, 3, 12), (16, 8, 8, 4, 14)])
def test_wcycle_matches_oracle(oracle_mod, precision, nelx, nely, nelz, L, seed):
    """mg.cycle = 1: the W-cycle of Implementation/WCycle.m as the preconditioner
    (SolveMG.m:31-37), against the oracle's restatement."""
    _say(f"wcycle {precision} {nelx}x{nely}x{nelz} L={L}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    s = _setup(prob, x, L, precision, cycle=1)
    b = np.random.default_rng(seed).standard_normal(prob.nfree)
    # FP64: explicit coarse inverse vs triangular solves; mixed: the FP32
    # cycle's rounding (eps32 = 6e-8) amplified through the two coarse visits
    assert _rel(s.vcycle(b), o.wcycle(b)) < (1e-8 if precision == "fp64" else 5e-5)
    u_or, it_or, _ = o.solve(prob.f, cycle=1)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    else:
        assert it <= it_or + 3
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


def test_mixed_solve_escalates_to_fp64_preconditioner(oracle_mod, monkeypatch):
    """A mixed-precision solve whose residual stops falling switches its
    preconditioner to the FP64 hierarchy mid-solve and restarts the
    directions. Forced here (a one-iteration window and an unreachable
    reduction factor): the result still meets the stop rule against the
    oracle's operator, and two forced solves are bitwise identical."""
    _say("mixed escalation")
    monkeypatch.setenv("TMG_STALL_WINDOW", "1")
    monkeypatch.setenv("TMG_STALL_FACTOR", "1e-300")
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 12)
    o = _oracle(oracle_mod, prob, x, 3)
    u_or, it_or, _ = o.solve(prob.f)
    s = _setup(prob, x, 3, "mixed")
    u, it, rr, rc = s.solve(prob.f)
    at, n = s.solve_stats()
    _say(f"  switched at iteration {at}, {it} iterations")
    assert n == 1 and 1 < at < it
    assert rc == 0 and rr <= 1e-8
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)
    u2, it2, _, _ = s.solve(prob.f)
    assert s.solve_stats() == (at, 2)
    assert it2 == it and np.array_equal(u2, u)
    # without the forcing the same handle's solves do not switch
    monkeypatch.delenv("TMG_STALL_WINDOW")
    monkeypatch.delenv("TMG_STALL_FACTOR")
    s2 = _setup(prob, x, 3, "mixed")
    s2.solve(prob.f)
    assert s2.solve_stats() == (0, 0)


def test_config1_uniform_density(oracle_mod):
    """48x24x24, rho = 0.3, 4 levels, line load; also the closed form
    u(rho) = u(1) / (Emin + (1 - Emin) rho^3) for uniform densities."""
    _say("config1")
    prob = P.cantilever(48, 24, 24)
    x = np.full(prob.nelem, 0.3)
    o = _oracle(oracle_mod, prob, x, 4)
    u_or, it_or, _ = o.solve(prob.f)
    for prec in ("fp64", "mixed"):
        s = _setup(prob, x, 4, prec)
        u, it, rr, rc = s.solve(prob.f)
        assert rc == 0
        _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)
        s1 = _setup(prob, np.ones(prob.nelem), 4, prec)
        u1, _, _, _ = s1.solve(prob.f)
        Emin = 1e-9
        assert _rel(u, u1 / (Emin + (1 - Emin) * 0.3 ** 3)) < 1e-6


def test_config2_fp64_and_mixed(oracle_mod):
    """96x48x48 (~699k DOFs), filtered random densities (seed 2), 4 levels."""
    _say("config2")
    prob = P.cantilever(96, 48, 48)
    x = P.random_density(96, 48, 48, 2)
    o = _oracle(oracle_mod, prob, x, 4)
    u_or, it_or, _ = o.solve(prob.f)
    for prec in ("fp64", "mixed"):
        s = _setup(prob, x, 4, prec)
        u, it, rr, rc = s.solve(prob.f)
        assert rc == 0
        _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_box_elements_match_oracle(oracle_mod, precision):
    """Non-cube elements (el, eh, ew = 1, 0.7, 1.3): the fine-level apply's
    spectral coefficients come from a general box Ke (every coefficient
    distinct), against the oracle's assembled operator and PCG."""
    _say(f"box elements {precision}")
    nelx, nely, nelz, L = 16, 8, 8, 3
    prob = P.from_rows(nelx, nely, nelz, [[0, 0, 0, nely, 0, nelz, 1, 1, 1]],
                       [[nelx, nelx, 0, 0, 0, nelz, 0, -1, 0]], l=16.0, h=5.6, w=10.4)
    assert (prob.el, prob.eh, prob.ew) == pytest.approx((1.0, 0.7, 1.3))
    x = P.random_density(nelx, nely, nelz, 21)
    o = _oracle(oracle_mod, prob, x, L)
    s = _setup(prob, x, L, precision)
    v = np.random.default_rng(21).standard_normal(prob.nfree)
    if precision == "fp64":
        assert _rel(s.apply(1, v), o.spmv(0, v)) < 1e-13
        b = np.random.default_rng(22).standard_normal(prob.nfree)
        assert _rel(s.smooth(1, b, v, 2), o.smooth(0, b, v, 2)) < 1e-12
    u_or, it_or, _ = o.solve(prob.f)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64":
        # ~120 iterations at this aspect ratio: the rounding differences of
        # the coarse solve (explicit inverse vs triangular solves) shift the
        # count by a few iterations, so it is compared to within 10%
        assert abs(it - it_or) <= 0.1 * it_or
    else:
        # the FP32 V-cycle is a weak preconditioner for these elements and
        # the count follows its rounding: measured 330 to ~800 iterations
        # (dense or spectral kernels, summation orders of the coarse sweeps)
        # against ~120 in FP64. Asserted: convergence within maxiterMG (rc 0
        # above) to the oracle's solution (below).
        _say(f"  mixed iterations {it} (oracle FP64 {it_or})")
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


@pytest.mark.parametrize("grid", [(30, 6, 6), (6, 30, 6), (40, 20, 12), (24, 70, 8)])
def test_spectral_apply_matches_dense_kernel(monkeypatch, grid):
    """The spectral element form of the fine-level apply against the dense
    24x24 element product (TMG_EBE_WHT=0) on the same inputs: the FP64
    apply, sweeps and V-cycle to rounding, and the mixed-precision V-cycle
    (FP32 fine-level sweeps) to FP32 rounding. The grids select each tile shape (TY = 8, 32, 16, 16) and
    leave partial tiles in x, y and z."""
    nelx, nely, nelz = grid
    _say(f"spectral vs dense {grid}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, 31)
    v = np.random.default_rng(31).standard_normal(prob.nfree)
    b = np.random.default_rng(32).standard_normal(prob.nfree)
    out = {}
    for tag, env in (("wht", "1"), ("dense", "0")):
        monkeypatch.setenv("TMG_EBE_WHT", env)
        s = _setup(prob, x, 2)
        out[tag] = (s.apply(1, v), s.smooth(1, b, v, 3))
        s.close()
        s = _setup(prob, x, 2, "mixed")
        out[tag + "32"] = s.vcycle(b)      # FP32 fine-level sweeps
        s.close()
        s = _setup(prob, x, 2)
        out[tag + "64"] = s.vcycle(b)
        s.close()
    assert _rel(out["wht"][0], out["dense"][0]) < 1e-14
    assert _rel(out["wht"][1], out["dense"][1]) < 1e-13
    assert _rel(out["wht64"], out["dense64"]) < 1e-13
    assert _rel(out["wht32"], out["dense32"]) < 1e-5
    assert _rel(out["wht32"], out["wht64"]) < 1e-5


@pytest.mark.parametrize("support_x", [0, 1])
def test_level1_matrix_free_matches_stencil(oracle_mod, monkeypatch, support_x):
    """Level 1 (the first Galerkin level) applied matrix-free in the
    multiresolution form (8 children's moduli, Galerkin integrands' table)
    against the assembled stencil (TMG_L1_STENCIL=1) and the oracle. A support
    on the odd plane x = 1 has fixed fine DOFs interpolating from free coarse
    DOFs: the operator is then not a sum over coarse elements and level 1 keeps
    the stencil; both cases must match the oracle."""
    nelx, nely, nelz, L = 16, 8, 8, 3
    _say(f"level-1 matrix-free, support x = {support_x}")
    prob = P.from_rows(nelx, nely, nelz, [[support_x, support_x, 0, nely, 0, nelz, 1, 1, 1]],
                       [[nelx, nelx, 0, 0, 0, nelz, 0, -1, 0]])
    x = P.random_density(nelx, nely, nelz, 51)
    o = _oracle(oracle_mod, prob, x, L)
    rng = np.random.default_rng(51)
    v1 = rng.standard_normal(o.nfree(1))
    b1 = rng.standard_normal(o.nfree(1))
    bf = rng.standard_normal(prob.nfree)
    out = {}
    for tag, env in (("free", "0"), ("stencil", "1")):
        monkeypatch.setenv("TMG_L1_STENCIL", env)
        s = _setup(prob, x, L)
        out[tag] = (s.apply(2, v1), s.smooth(2, b1, v1, 2), s.vcycle(bf))
        s.close()
        s = _setup(prob, x, L, "mixed")
        out[tag + "32"] = s.vcycle(bf)
        s.close()
    assert _rel(out["free"][0], o.spmv(1, v1)) < 1e-13
    assert _rel(out["free"][1], o.smooth(1, b1, v1, 2)) < 1e-12
    for i, tol in enumerate((1e-13, 1e-12, 1e-10)):
        assert _rel(out["free"][i], out["stencil"][i]) < tol, i
    assert _rel(out["free32"], out["stencil32"]) < 1e-5
    assert _rel(out["free32"], out["free"][2]) < 1e-5


def test_bitwise_repeatable():
    _say("repeatable")
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 5)
    for prec in ("fp64", "mixed"):
        s = _setup(prob, x, 3, prec)
        u1, it1, _, _ = s.solve(prob.f)
        u2, it2, _, _ = s.solve(prob.f)
        s2 = _setup(prob, x, 3, prec)
        u3, it3, _, _ = s2.solve(prob.f)
        assert it1 == it2 == it3
        assert np.array_equal(u1, u2) and np.array_equal(u1, u3)


def test_warm_start_and_not_converged(oracle_mod):
    """Warm start x0 (pcgp.opu0, Implementation/StrSLP.m:329-333) and the
    silent stop at maxiterMG (Implementation/SolveMG.m:30)."""
    _say("warm start")
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 6)
    s = _setup(prob, x, 3)
    u, it, rr, rc = s.solve(prob.f)
    u2, it2, rr2, rc2 = s.solve(prob.f, x0=u)
    assert rc2 == 0 and it2 <= 2
    s1 = _setup(prob, x, 3, maxiter=1)
    _, it1, rr1, rc1 = s1.solve(prob.f)
    assert rc1 == 1 and it1 == 2 and rr1 > 1e-8
    s0 = _setup(prob, x, 3, tol=0.0, maxiter=3)
    _, it0, _, rc0 = s0.solve(prob.f)
    assert rc0 == 1 and it0 == 4


def test_grad_and_filter_match_oracle(oracle_mod):
    _say("grad/filter")
    import tmg
    nelx, nely, nelz = 16, 8, 8
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, 7)
    o = _oracle(oracle_mod, prob, x, 3)
    u_or, _, _ = o.solve(prob.f)
    ufull = prob.full(u_or)
    s = _setup(prob, x, 3)
    for opf, rmin in ((2, 1.5), (1, 2.2), (2, 2.5)):
        W, w = oracle_mod.filter_matrix(nelx, nely, nelz, rmin, opf)
        f = tmg.Filter(nelx, nely, nelz, rmin, opf)
        assert _rel(f.rowsum(), w) < 1e-15
        v = np.random.default_rng(opf).random(prob.nelem)
        assert _rel(f.apply(v), (W @ v) / w) < 1e-14
        assert _rel(f.apply(v, trans=True), W.T @ (v / w)) < 1e-14
        g0 = o.grad0(ufull, x)
        assert _rel(s.grad(ufull, x), g0) < 1e-12
        assert _rel(s.grad(ufull, x, filt=f), W.T @ (g0 / w)) < 1e-12
    fixed = np.zeros(prob.nelem, dtype=np.uint8)
    fixed[::7] = 1
    W, w = oracle_mod.filter_matrix(nelx, nely, nelz, 1.5, 2, fixed=fixed)
    f = tmg.Filter(nelx, nely, nelz, 1.5, 2, fixed=fixed)
    v = np.random.default_rng(9).random(prob.nelem)
    assert _rel(f.apply(v), (W @ v) / w) < 1e-14
    freedens = np.ones(prob.nelem, dtype=np.uint8)
    freedens[::5] = 0
    assert _rel(s.grad(ufull, x, freedens=freedens), o.grad0(ufull, x, freedens)) < 1e-12


def test_error_paths():
    _say("error paths")
    import tmg
    prob = P.cantilever(8, 4, 4)
    s = tmg.Solver(prob, 3)
    with pytest.raises(tmg.TmgError) as e:
        s.solve(prob.f)   # no operator yet
    assert e.value.code == -6
    with pytest.raises(tmg.TmgError):
        s.update_density(-np.ones(prob.nelem))
    s.update_density(np.full(prob.nelem, 0.5))
    with pytest.raises(tmg.TmgError):
        s.apply(9, np.zeros(3))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multiresolution_matches_oracle(oracle_mod, precision):
    """mr.n = 2: operator, Galerkin hierarchy, solve and density-mesh gradient
    against the oracle's multiresolution assembly (Implementation/GlobalStiffMatrix.m:19-62,
    Implementation/Grad.m:24-49)."""
    _say(f"multiresolution {precision}")
    import tmg
    nelx, nely, nelz, L = 8, 4, 4, 3
    prob = P.cantilever(nelx, nely, nelz)
    xd = P.random_density(2 * nelx, 2 * nely, 2 * nelz, 21)
    o = oracle_mod.Oracle(prob)
    o.assemble_mr(2, xd)
    o.setup_gmg(L, 0.6)
    s = tmg.Solver(prob, L, precision=precision, mr=2)
    s.update_density(xd)
    rng = np.random.default_rng(3)
    for lev in range(1, L + 1):
        v = rng.standard_normal(o.nfree(lev - 1))
        assert _rel(s.apply(lev, v), o.spmv(lev - 1, v)) < 1e-13
    u_or, it_or, _ = o.solve(prob.f)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    ufull = prob.full(u_or)
    assert _rel(s.grad(ufull, xd), o.grad0_mr(2, ufull, xd)) < 1e-12
    f = tmg.Filter(2 * nelx, 2 * nely, 2 * nelz, 1.5, 2, el=0.5, eh=0.5, ew=0.5)
    W, w = oracle_mod.filter_matrix(2 * nelx, 2 * nely, 2 * nelz, 1.5, 2, el=0.5, eh=0.5, ew=0.5)
    g0 = o.grad0_mr(2, ufull, xd)
    assert _rel(s.grad(ufull, xd, filt=f), W.T @ (g0 / w)) < 1e-12


@pytest.mark.parametrize("grid", [(8, 4, 4), (20, 6, 10)])
def test_multiresolution_spectral_matches_stencil(monkeypatch, grid):
    """The matrix-free multiresolution apply (spectral element form with the
    sub-element moduli's transform) against the assembled fine-level block
    stencil (TMG_MR_STENCIL=1) on the same densities: FP64 apply and sweeps to
    rounding, and the mixed-precision V-cycle to FP32 rounding."""
    import tmg
    nelx, nely, nelz = grid
    _say(f"multiresolution spectral vs stencil {grid}")
    prob = P.cantilever(nelx, nely, nelz)
    xd = P.random_density(2 * nelx, 2 * nely, 2 * nelz, 41)
    v = np.random.default_rng(41).standard_normal(prob.nfree)
    b = np.random.default_rng(42).standard_normal(prob.nfree)
    out = {}
    for tag, env in (("wht", "0"), ("stencil", "1")):
        monkeypatch.setenv("TMG_MR_STENCIL", env)
        for prec in ("fp64", "mixed"):
            s = tmg.Solver(prob, 2, precision=prec, mr=2)
            s.update_density(xd)
            out[tag, prec] = (s.apply(1, v), s.smooth(1, b, v, 2), s.vcycle(b))
            s.close()
    for i, tol in enumerate((1e-13, 1e-12, 1e-10)):
        assert _rel(out["wht", "fp64"][i], out["stencil", "fp64"][i]) < tol, i
    assert _rel(out["wht", "mixed"][2], out["stencil", "mixed"][2]) < 1e-5
    assert _rel(out["wht", "mixed"][2], out["wht", "fp64"][2]) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz,L,seed,run", [(16, 8, 8, 3, 21, 8), (24, 12, 12, 3, 22, 8),
                                                       (24, 12, 12, 3, 25, 2), (32, 16, 16, 3, 26, 3)])
def test_symmetric_coarse_solve_matches_full_and_oracle(oracle_mod, monkeypatch, precision, nelx, nely, nelz, L,
                                                        seed, run):
    """The coarse solve that reads only the upper block triangle of the
    symmetric inverse (k_coarse_symv_run/sum, the default from 16,384 coarse
    DOFs) forced on small coarse systems whose size is not a multiple of the
    64-row block (108, 288 and 600 DOFs), with runs of 8, 2 and 3 tiles (so
    that block rows have several runs, a partial last run, and a last block
    row and column that end inside a tile): the V-cycle equals the
    full-matrix solve's to rounding, and the PCG solve meets the oracle."""
    _say(f"symmetric coarse solve {precision} {nelx}x{nely}x{nelz} L={L} run={run}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    monkeypatch.setenv("TMG_COARSE_SYMV", "0")
    s0 = _setup(prob, x, L, precision)
    monkeypatch.setenv("TMG_COARSE_SYMV", "1")
    monkeypatch.setenv("TMG_CS_RUN", str(run))
    s1 = _setup(prob, x, L, precision)
    nc = s1.level_info(L)[1].size
    assert nc % 64 != 0
    b = np.random.default_rng(seed).standard_normal(prob.nfree)
    v0, v1 = s0.vcycle(b), s1.vcycle(b)
    assert _rel(v1, v0) < (1e-13 if precision == "fp64" else 1e-6)
    if precision == "fp64":
        assert _rel(v1, o.vcycle(b)) < 1e-8
    u_or, it_or, _ = o.solve(prob.f)
    u, it, rr, rc = s1.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64" and abs(it - it_or) > 1:
        # the recursive residual of this problem plateaus near the tolerance
        # (oracle: 1.64e-8 after 29 iterations, 1.76e-8 after 30, 9.7e-9
        # after 31), so rounding may move the stop by two: the oracle's own
        # residual at the iteration where the device stopped is within 2x tol
        _, _, rr_or = o.solve(prob.f, maxiter=it - 1)
        _say(f"  it {it}, oracle {it_or}; oracle relres after {it - 1} iterations {rr_or:.3e}")
        assert abs(it - it_or) <= 2 and rr_or <= 2e-8
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz,L,seed,leaf", [(16, 8, 8, 3, 31, 4), (24, 12, 12, 3, 32, 8),
                                                        (32, 16, 16, 3, 33, 16), (64, 32, 32, 3, 34, 64)])
def test_nested_dissection_coarse_solve_matches_dense_and_oracle(oracle_mod, monkeypatch, precision, nelx, nely,
                                                                 nelz, L, seed, leaf):
    """The coarse solve by nested dissection and multifrontal Cholesky (the
    default from 16,384 coarse DOFs) forced on small coarse grids, with leaves
    of 4 ... 64 nodes so that the dissection tree has 2 to 6 levels: the
    V-cycle equals the dense-inverse solve's to rounding, and the PCG solve
    meets the oracle."""
    _say(f"nested-dissection coarse solve {precision} {nelx}x{nely}x{nelz} L={L} leaf={leaf}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    monkeypatch.setenv("TMG_COARSE_ND", "0")
    monkeypatch.setenv("TMG_COARSE_F32", "0")
    s0 = _setup(prob, x, L, precision)
    monkeypatch.setenv("TMG_COARSE_ND", "1")
    monkeypatch.setenv("TMG_ND_LEAF", str(leaf))
    s1 = _setup(prob, x, L, precision)
    b = np.random.default_rng(seed).standard_normal(prob.nfree)
    v0, v1 = s0.vcycle(b), s1.vcycle(b)
    assert _rel(v1, v0) < (1e-12 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        assert _rel(v1, o.vcycle(b)) < 1e-8
    u_or, it_or, _ = o.solve(prob.f)
    u, it, rr, rc = s1.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    else:
        # the FP32 cycle's iteration count moves with rounding; the dense
        # solve's count on the same problem is the reference for it
        _, it0, _, _ = s0.solve(prob.f)
        _say(f"  mixed: {it} iterations, dense coarse solve {it0}, oracle (FP64) {it_or}")
        assert abs(it - it0) <= 2 and it <= 1.5 * it_or + 3
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_consumer_side_pcg_reductions_are_bitwise_the_finalize_kernels(monkeypatch, precision):
    """rho/beta and alpha reduced by every block of k_dir_xupdate and
    k_r_update (the default) give bitwise the iterates of the one-block
    k_finalize launches (TMG_FIN_CONSUMER=0): same reduction order."""
    prob = P.cantilever(32, 16, 16)
    x = P.random_density(32, 16, 16, 41)
    out = {}
    for env in ("0", "1"):
        monkeypatch.setenv("TMG_FIN_CONSUMER", env)
        s = _setup(prob, x, 4, precision)
        out[env] = s.solve(prob.f)
        s.close()
    (u0, it0, rr0, rc0), (u1, it1, rr1, rc1) = out["0"], out["1"]
    assert rc0 == 0 and rc1 == 0 and it0 == it1 and rr0 == rr1
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("cycle", [0, 1])
def test_restriction_with_first_sweep_is_bitwise_the_separate_sweep(monkeypatch, precision, cycle):
    """The restriction that also forms the next level's first Jacobi sweep
    from zero (the default) gives bitwise the cycle and the solve of the
    separate k_jacobi_zero launch (TMG_RESTRICT_JZ=0). The small levels'
    fused sweeps are switched off (TMG_FUSED_MAXNODES=0) so that every
    level between the finest and the coarsest takes that path."""
    prob = P.cantilever(32, 16, 16)
    x = P.random_density(32, 16, 16, 47)
    b = np.random.default_rng(47).standard_normal(prob.nfree)
    monkeypatch.setenv("TMG_FUSED_MAXNODES", "0")
    out = {}
    for env in ("0", "1"):
        monkeypatch.setenv("TMG_RESTRICT_JZ", env)
        s = _setup(prob, x, 4, precision, cycle=cycle)
        out[env] = (s.vcycle(b), s.solve(prob.f))
        s.close()
    (v0, (u0, it0, rr0, rc0)), (v1, (u1, it1, rr1, rc1)) = out["0"], out["1"]
    assert np.array_equal(v0, v1)
    assert rc0 == 0 and rc1 == 0 and it0 == it1 and rr0 == rr1
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_prolongation_per_coarse_node_is_bitwise_the_per_fine_node_form(monkeypatch, precision):
    """k_prolong_add4 (one thread per coarse node and fine plane, the default)
    gives bitwise the values of k_prolong_add (one thread per fine node,
    TMG_PROLONG4=0): same terms, weights and order. Checked on every level's
    prolongation and on whole solves (W-cycle: each level prolongs twice)."""
    prob = P.cantilever(32, 16, 16)
    x = P.random_density(32, 16, 16, 43)
    out = {}
    for env in ("0", "1"):
        monkeypatch.setenv("TMG_PROLONG4", env)
        s = _setup(prob, x, 4, precision, cycle=1)
        pro = [s.prolong(lev, np.random.default_rng(lev).standard_normal(s.nfree_level(lev + 1)))
               for lev in range(1, 4)]
        out[env] = (pro, s.solve(prob.f))
        s.close()
    (p0, (u0, it0, rr0, rc0)), (p1, (u1, it1, rr1, rc1)) = out["0"], out["1"]
    for a, b in zip(p0, p1):
        assert np.array_equal(a, b)
    assert rc0 == 0 and rc1 == 0 and it0 == it1 and rr0 == rr1
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("leaf,cycle", [(8, 0), (27, 0), (8, 1), (27, 2)])
def test_nested_dissection_with_roller_supports(oracle_mod, monkeypatch, leaf, cycle):
    """Nested dissection on a quarter-symmetry beam whose supports fix single
    displacement components (symmetry planes x = 0 and z = 0, a roller line):
    coarse nodes with one or two free DOFs, under the V, W and FMV cycles.
    The preconditioner equals the dense solve's to rounding; the PCG solve
    meets the oracle's with the same cycle."""
    nelx, nely, nelz, L = 32, 16, 16, 3
    _say(f"nested dissection, roller supports, leaf={leaf}, cycle={cycle}")
    supp = [[0, 0, 0, nely, 0, nelz, 1, 0, 0], [0, nelx, 0, nely, 0, 0, 0, 0, 1], [nelx, nelx, 0, 0, 0, nelz, 0, 1, 0]]
    prob = P.from_rows(nelx, nely, nelz, supp, [[0, 0, nely, nely, 0, nelz, 0, -1, 0]])
    x = P.random_density(nelx, nely, nelz, 35)
    o = _oracle(oracle_mod, prob, x, L)
    monkeypatch.setenv("TMG_COARSE_ND", "0")
    s0 = _setup(prob, x, L, "fp64", cycle=cycle)
    monkeypatch.setenv("TMG_COARSE_ND", "1")
    monkeypatch.setenv("TMG_ND_LEAF", str(leaf))
    s1 = _setup(prob, x, L, "fp64", cycle=cycle)
    b = np.random.default_rng(35).standard_normal(prob.nfree)
    assert _rel(s1.vcycle(b), s0.vcycle(b)) < 1e-12
    u_or, it_or, _ = o.solve(prob.f, cycle=cycle)
    u, it, rr, rc = s1.solve(prob.f)
    assert rc == 0 and rr <= 1e-8 and abs(it - it_or) <= 1
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz,L,seed", [(16, 8, 8, 3, 23), (16, 8, 8, 4, 24)])
def test_fmv_cycle_matches_oracle(oracle_mod, precision, nelx, nely, nelz, L, seed):
    """mg.cycle = 2: the FMV cycle of Implementation/FMVCycle.m as the
    preconditioner (SolveMG.m:31-37), against the oracle's restatement."""
    _say(f"fmv {precision} {nelx}x{nely}x{nelz} L={L}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed)
    o = _oracle(oracle_mod, prob, x, L)
    s = _setup(prob, x, L, precision, cycle=2)
    b = np.random.default_rng(seed).standard_normal(prob.nfree)
    assert _rel(s.vcycle(b), o.fmvcycle(b)) < (1e-8 if precision == "fp64" else 5e-5)
    u_or, it_or, _ = o.solve(prob.f, cycle=2)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0 and rr <= 1e-8
    if precision == "fp64":
        assert abs(it - it_or) <= 1
    else:
        assert it <= it_or + 3
    _check_solution(oracle_mod, o, prob, u, it, it_or, u_or, 1e-6)


def test_time_in_solve_counts_the_iterations_run():
    """tmg_time_in_solve times each launch of the FP64 apply and of the fine
    post-sweep in one eagerly launched solve: one of each per PCG iteration,
    none from the iterations skipped past convergence; the timed solve gives
    the same solution as the graph-replayed one."""
    _say("time in solve")
    import tmg
    prob = P.cantilever(32, 16, 16)
    x = P.random_density(32, 16, 16, 7)
    for prec in ("fp64", "mixed"):
        s = _setup(prob, x, 4, prec)
        u, it, _, _ = s.solve(prob.f)
        ms0, n0 = s.time_in_solve(0)
        ms8, n8 = s.time_in_solve(8)
        assert n0 == it - 1 and n8 == it - 1 and ms0 > 0 and ms8 > 0
        with pytest.raises(tmg.TmgError):
            s.time_in_solve(3)
        u2, it2, _, _ = s.solve(prob.f)
        assert it2 == it and np.array_equal(u, u2)
        s.close()
