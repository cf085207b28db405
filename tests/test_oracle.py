"""Pins of the FP64 oracle against the reference's own known answers and
against what the mathematics fixes (closed forms, invariants, independent
restatements). CPU only."""
import hashlib
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from tmg import problem as P

HERE = os.path.dirname(os.path.abspath(__file__))
PERM2 = [1, 2, 3, 7, 8, 9, 10, 11, 12, 4, 5, 6, 13, 14, 15, 19, 20, 21, 22, 23, 24, 16, 17, 18]


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------------------
# element stiffness
# ---------------------------------------------------------------------------
def _analytic_rows(el, eh, ew, v):
    """Closed-form entries of rows 1 and 2 of the reference's element matrix
    before its perm2 reordering (Implementation/ElemStiffMatrix.m:7-47)."""
    a, b, c = el / 2, eh / 2, ew / 2
    al2, al3 = a * a * c * c * (-1 + v), b * b * c * c * (-1 + v)
    be1, be2, be3 = a * a * b * b * (1 - 2 * v), a * a * c * c * (1 - 2 * v), b * b * c * c * (1 - 2 * v)
    g, d = 9 * a * b * c, 1.0 / ((1 + v) * (1 - 2 * v))
    k = {}
    k[1, 1] = d * (-2 * al3 + be1 + be2) / g
    k[1, 2] = d * c / 12
    k[1, 3] = d * b / 12
    k[1, 4] = d * (4 * al3 + be1 + be2) / (2 * g)
    k[1, 5] = d * (4 * c * v - c) / 12
    k[1, 6] = d * (4 * b * v - b) / 12
    k[1, 7] = d * (4 * al3 + be1 - 2 * be2) / (4 * g)
    k[1, 8] = -k[1, 2]
    k[1, 9] = k[1, 6] / 2
    k[1, 10] = d * (-2 * al3 + be1 - 2 * be2) / (2 * g)
    k[1, 11] = -k[1, 5]
    k[1, 12] = k[1, 3] / 2
    k[1, 13] = d * (-2 * al3 - 2 * be1 + be2) / (2 * g)
    k[1, 14] = k[1, 2] / 2
    k[1, 15] = -k[1, 6]
    k[1, 16] = d * (4 * al3 - 2 * be1 + be2) / (4 * g)
    k[1, 17] = k[1, 5] / 2
    k[1, 18] = -k[1, 3]
    k[1, 19] = -k[1, 1] / 4
    k[1, 20] = -k[1, 14]
    k[1, 21] = -k[1, 12]
    k[1, 22] = d * (-al3 - be1 - be2) / (2 * g)
    k[1, 23] = -k[1, 17]
    k[1, 24] = -k[1, 9]
    k[2, 2] = d * (be1 - 2 * al2 + be3) / g
    k[2, 3] = d * a / 12
    k[2, 5] = d * (be1 - 2 * al2 - 2 * be3) / (2 * g)
    k[2, 8] = d * (be1 + 4 * al2 - 2 * be3) / (4 * g)
    k[2, 9] = d * (4 * a * v - a) / 24
    k[2, 11] = d * (be1 + 4 * al2 + be3) / (2 * g)
    k[2, 14] = d * (be3 - 2 * al2 - 2 * be1) / (2 * g)
    return k


@pytest.mark.parametrize("el,eh,ew,nu", [(1, 1, 1, 0.3), (1.0, 0.7, 1.3, 0.25), (0.5, 0.5, 0.5, 0.3)])
def test_ke_matches_reference_closed_form(oracle_mod, el, eh, ew, nu):
    K = oracle_mod.elem_stiff(el, eh, ew, nu)
    for (i, j), val in _analytic_rows(el, eh, ew, nu).items():
        # k(perm2, perm2) = k  (ElemStiffMatrix.m:331-332)
        got = K[PERM2[i - 1] - 1, PERM2[j - 1] - 1]
        assert abs(got - val) <= 1e-13 * max(1.0, abs(val)), (i, j, got, val)


def test_ke_symmetry_rigid_modes_and_scaling(oracle_mod):
    K = oracle_mod.elem_stiff(1.0, 0.8, 1.2, 0.3)
    assert np.array_equal(K, K.T)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12 * ev.max()) == 6 and ev.min() > -1e-12
    # rigid body modes: 3 translations, 3 rotations about the element centre
    X = np.array([[(a >> 1) & 1, a & 1, a >> 2] for a in range(8)], float) * [1.0, 0.8, 1.2]
    modes = []
    for c in range(3):
        m = np.zeros(24)
        m[c::3] = 1
        modes.append(m)
    for (i, j) in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros(24)
        m[i::3] = -X[:, j]
        m[j::3] = X[:, i]
        modes.append(m)
    for m in modes:
        assert np.linalg.norm(K @ m) < 1e-13 * np.linalg.norm(K) * np.linalg.norm(m)
    # Ke(h el) = h Ke(el) for 3-D Q1 elasticity
    assert _rel(oracle_mod.elem_stiff(2.0, 1.6, 2.4, 0.3), 2.0 * K) < 1e-14


def test_mr_integrands_sum_to_ke(oracle_mod):
    K = oracle_mod.elem_stiff(1, 1, 1, 0.3)
    for n in (1, 2, 3):
        k = oracle_mod.elem_stiff_mr(n, 1, 1, 1, 0.3)
        assert _rel(k.sum(axis=0), K) < 1e-14
        for ki in k:
            assert np.allclose(ki, ki.T, atol=0)


# ---------------------------------------------------------------------------
# problem set-up against the reference's own problem files
# ---------------------------------------------------------------------------
def _fixtures():
    return json.load(open(os.path.join(HERE, "golden", "setup_fixtures.json")))


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i8").tobytes()).hexdigest()


@pytest.mark.parametrize("fx", _fixtures(), ids=lambda f: f["file"])
def test_problem_generator_matches_reference_files(fx):
    """Supports, point / line / area loads and fixed-density rows of the
    reference's own problem files (24 of them keep their rows) rebuild their
    stored str.freedofs, str.f and str.fixedDens / fixedDensVal."""
    prob = P.from_rows(fx["nelx"], fx["nely"], fx["nelz"], fx["supp"], fx["loads"], l=fx["l"], h=fx["h"], w=fx["w"],
                       fix=fx.get("fix"))
    assert prob.nfree == fx["nfree"]
    assert _sha(prob.freedofs) == fx["freedofs_sha256"]
    nz = np.nonzero(prob.f)[0]
    np.testing.assert_array_equal(nz + 1, fx["f_nonzero_index1"])
    np.testing.assert_allclose(prob.f[nz], fx["f_nonzero_value"], rtol=1e-15, atol=0)
    assert prob.fixed_dens.size == fx["fixed_dens_count"] and _sha(prob.fixed_dens) == fx["fixed_dens_sha256"]
    assert int(np.sum(prob.fixed_dens_val == 1)) == fx["fixed_dens_solid"]


def test_fixture_library_covers_area_loads_and_fixed_regions():
    fxs = _fixtures()
    area = [f["file"] for f in fxs if any(sum(r[2 * k + 1] > r[2 * k] for k in range(3)) == 2 for r in f["loads"])]
    fixed = [f["file"] for f in fxs if f.get("fix")]
    assert len(fxs) == 24 and len(area) >= 2 and len(fixed) >= 4


def test_area_load_shares():
    """A load on an area spreads the line rule's shares in both directions
    (Implementation/EnterData.m:350-782): a unit pressure on the face y = 0 of
    a 4x2x3 block gives each node the product of its x and z shares (1/2 at
    the ends, 1 inside) and totals the face area; loads on the faces z = 0
    (xy plane) and x = 4 (yz plane) total theirs; a patch inside one node's
    tributary area puts its whole area on that node."""
    supp = [[0, 4, 2, 2, 0, 3, 1, 1, 1]]            # top face y = 2: no loaded node is supported
    g = P.from_rows(4, 2, 3, supp, [[0, 4, 0, 0, 0, 3, 0, -1, 0]])
    f = g.full(g.f).reshape(-1, 3)
    nx, ny = 5, 3
    wx, wz = [0.5, 1, 1, 1, 0.5], [0.5, 1, 1, 0.5]
    for kz in range(4):
        for kx in range(5):
            assert f[kz * nx * ny + kx * ny, 1] == -wx[kx] * wz[kz]
    assert abs(g.f.sum() + 12.0) < 1e-12 and np.count_nonzero(g.f) == 20
    xy = P.from_rows(4, 2, 3, [[0, 4, 0, 2, 3, 3, 1, 1, 1]], [[0, 4, 0, 2, 0, 0, 0, 0, 2.0]])
    assert abs(xy.f.sum() - 2.0 * 8) < 1e-12 and np.count_nonzero(xy.f) == 15
    yz = P.from_rows(4, 2, 3, supp, [[4, 4, 0, 2, 0, 3, 1.0, 0, 0]])
    assert abs(yz.f.sum() - (2 * 3 - 0.5 * 3)) < 1e-12          # the y = 2 edge is supported
    h = P.from_rows(4, 2, 3, supp, [[1.1, 1.3, 0, 0, 2.2, 2.4, 0, -1, 0]])
    nzv = h.f[np.nonzero(h.f)]
    assert nzv.size == 1 and nzv[0] == -((1.3 - 1.1) * (2.4 - 2.2))
    with pytest.raises(ValueError):
        P.from_rows(4, 2, 3, supp, [[0, 4, 0, 2, 0, 3, 0, -1, 0]])   # volumetric


def test_fixed_regions_element_layers():
    """Void / solid regions by coordinates become element layers as
    Implementation/EnterData.m:838-843 rounds them, clamped to the grid."""
    rows, dens, val = P.fixed_regions(8, 4, 4, 1.0, 1.0, 1.0, [[0, 2, 0, 4, 0, 4, 0], [7.6, 9, 1.4, 2.6, 0, 1, 1]])
    assert rows.tolist() == [[1, 2, 1, 4, 1, 4, 0], [8, 8, 1, 3, 1, 1, 1]]
    assert dens.size == 2 * 4 * 4 + 3 and int(val.sum()) == 3
    p = P.from_rows(8, 4, 4, [[0, 0, 0, 4, 0, 4, 1, 1, 1]], [[8, 8, 0, 0, 0, 4, 0, -1, 0]], fix=rows)
    assert np.array_equal(p.fixed_dens, dens) and np.array_equal(p.fixed_dens_val, val)
    assert p.free_dens_mask().sum() == p.nelem - dens.size


def test_line_load_shares():
    prob = P.cantilever(48, 24, 24)
    f = prob.full(prob.f).reshape(-1, 3)
    ny, nx = 25, 49
    nodes = [iz * nx * ny + 48 * ny + 0 for iz in range(25)]
    fy = f[nodes, 1]
    assert fy[0] == -0.5 and fy[-1] == -0.5 and np.all(fy[1:-1] == -1.0)
    assert np.count_nonzero(f) == 25 and f.sum() == -24.0


def test_choose_ngrids_rules():
    assert P.choose_ngrids(48, 24, 24, 4) == (4, 0, True)
    assert P.choose_ngrids(80, 40, 40, 5) == (4, 0, True)
    assert P.choose_ngrids(112, 56, 56, 5) == (4, 0, True)
    assert P.choose_ngrids(32, 16, 16, 5) == (5, 0, True)
    assert P.choose_ngrids(240, 120, 120, 3)[0] == 4          # nmaxchol forces a 4th level
    assert P.choose_ngrids(7, 4, 4, 3) == (3, 0, False)        # no GMG possible
    assert P.choose_ngrids(30, 30, 30, 2, nmaxchol=100) == (2, 1, True)   # cannot add a level


# ---------------------------------------------------------------------------
# operator, prolongation, Galerkin
# ---------------------------------------------------------------------------
def _assemble_numpy(prob, x, Ke, Emin=1e-9, p=3.0):
    """Independent scipy assembly of K(free, free) from Ke and the DOF map."""
    nelx, nely, nelz = prob.nelx, prob.nely, prob.nelz
    nx, ny = nelx + 1, nely + 1
    E = Emin + (1 - Emin) * x ** p
    rows, cols, vals = [], [], []
    e = 0
    for ez in range(nelz):
        for ex in range(nelx):
            for ey in range(nely):
                dofs = []
                for a in range(8):
                    dy, dx, dz = a & 1, (a >> 1) & 1, a >> 2
                    n = (ez + dz) * nx * ny + (ex + dx) * ny + (ey + dy)
                    dofs += [3 * n, 3 * n + 1, 3 * n + 2]
                dofs = np.array(dofs)
                rows.append(np.repeat(dofs, 24))
                cols.append(np.tile(dofs, 24))
                vals.append((E[e] * Ke).ravel())
                e += 1
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(prob.ndof, prob.ndof))
    fd = prob.freedofs - 1
    return K[fd][:, fd]


def test_assembled_operator_matches_independent_assembly(oracle_mod):
    prob = P.cantilever(6, 4, 2)
    x = P.random_density(6, 4, 2, 1)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    K = o.csr("A", 0)
    Kn = _assemble_numpy(prob, x, oracle_mod.elem_stiff())
    assert abs(K - Kn).max() < 1e-15 * abs(Kn).max()
    v = np.random.default_rng(0).standard_normal(prob.nfree)
    assert _rel(o.kelemprod(v), Kn @ v) < 1e-14
    # uniform density: K(rho) = E(rho) K(1)
    o.assemble(np.full(prob.nelem, 0.4))
    K4 = o.csr("A", 0)
    o.assemble(np.ones(prob.nelem))
    K1 = o.csr("A", 0)
    assert abs(K4 - (1e-9 + (1 - 1e-9) * 0.4 ** 3) * K1).max() < 1e-15 * abs(K1).max()


def _prolongation_transcribed(nxf, nyf, nzf, nxc, nyc, nzc, freedofsf1):
    """Loop-by-loop transcription of Implementation/Prolongation.m (1-based)."""
    w = [1, 0.5, 0.25, 0.125]
    nc = nxc * nyc * nzc
    I, J, V = [], [], []
    freedofsc = np.zeros(3 * nc, dtype=np.int64)
    aux = np.zeros(3 * nc, dtype=np.int64)
    for zc in range(1, nzc + 1):
        for xc in range(1, nxc + 1):
            for yc in range(1, nyc + 1):
                col = (zc - 1) * nxc * nyc + (xc - 1) * nyc + yc
                xf, yf, zf = 2 * xc - 1, 2 * yc - 1, 2 * zc - 1
                for z in range(max(zf - 1, 1), min(zf + 1, nzf) + 1):
                    for x in range(max(xf - 1, 1), min(xf + 1, nxf) + 1):
                        for y in range(max(yf - 1, 1), min(yf + 1, nyf) + 1):
                            row = (z - 1) * nxf * nyf + (x - 1) * nyf + y
                            val = (xf - x) ** 2 + (yf - y) ** 2 + (zf - z) ** 2 + 1
                            for k in (2, 1, 0):
                                I.append(3 * row - k)
                                J.append(3 * col - k)
                                V.append(w[val - 1])
                            if val == 1:
                                for k in (2, 1, 0):
                                    freedofsc[3 * col - k - 1] = 3 * col - k
                                    aux[3 * col - k - 1] = 3 * row - k
    Pm = sp.csr_matrix((V, (np.array(I) - 1, np.array(J) - 1)), shape=(3 * nxf * nyf * nzf, 3 * nc))
    freedofsc = freedofsc[np.isin(aux, freedofsf1)]
    return Pm[freedofsf1 - 1][:, freedofsc - 1], freedofsc


def test_prolongation_matches_transcription_and_trilinear_fields(oracle_mod):
    prob = P.from_rows(4, 2, 2, [[0, 0, 0, 2, 0, 2, 1, 0, 1]], [[4, 4, 0, 0, 1, 1, 0, -1, 0]])
    o = oracle_mod.Oracle(prob)
    o.assemble(np.full(prob.nelem, 0.5))
    o.setup_gmg(2, 0.6, allow_singular=True)   # y translation is free here
    Pm = o.csr("P", 0)
    Pt, fdc = _prolongation_transcribed(5, 3, 3, 3, 2, 2, prob.freedofs)
    np.testing.assert_array_equal(o.level_free(1) + 1, fdc)
    assert abs(Pm - Pt).max() == 0.0
    # with no supports P reproduces trilinear coarse fields exactly
    free = P.from_rows(8, 4, 4, np.zeros((0, 9)), [[8, 8, 0, 0, 2, 2, 0, -1, 0]])
    o2 = oracle_mod.Oracle(free)
    o2.assemble(np.full(free.nelem, 0.5))
    o2.setup_gmg(2, 0.6, allow_singular=True)   # no supports: K is singular
    P2 = o2.csr("P", 0)
    cx, cy, cz = np.meshgrid(np.arange(5), np.arange(3), np.arange(3), indexing="ij")
    fxg, fyg, fzg = np.meshgrid(np.arange(9), np.arange(5), np.arange(5), indexing="ij")

    def field(gx, gy, gz, scale):
        # node order z, x, y (y fastest); linear field in physical coords
        X, Y, Z = (np.transpose(g, (2, 0, 1)).ravel() * scale for g in (gx, gy, gz))
        v = np.zeros(3 * X.size)
        v[0::3], v[1::3], v[2::3] = 1 + 2 * X - Y + 0.5 * Z, X * 0.3 - Z, 2.0 + Y
        return v
    assert _rel(P2 @ field(cx, cy, cz, 2.0), field(fxg, fyg, fzg, 1.0)) < 1e-15


def test_galerkin_identity_and_scaling(oracle_mod):
    prob = P.cantilever(8, 4, 4)
    x = P.random_density(8, 4, 4, 3)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(3, 0.6)
    rng = np.random.default_rng(1)
    for lev in (0, 1):
        Pm = o.csr("P", lev)
        A, Ac = o.csr("A", lev), o.csr("A", lev + 1)
        xc = rng.standard_normal(Ac.shape[0])
        lhs = xc @ (Ac @ xc)
        rhs = (Pm @ xc) @ (A @ (Pm @ xc))
        assert abs(lhs - rhs) < 1e-12 * abs(rhs)
        assert _rel(Ac.toarray(), Ac.toarray().T) < 1e-14


# ---------------------------------------------------------------------------
# smoother, V-cycle, PCG
# ---------------------------------------------------------------------------
def test_two_level_cycle_closed_form(oracle_mod):
    prob = P.cantilever(4, 2, 2)
    x = P.random_density(4, 2, 2, 4)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(2, 0.6)
    A = o.csr("A", 0).toarray()
    Pm = o.csr("P", 0).toarray()
    Ac = Pm.T @ A @ Pm
    Minv = np.diag(0.6 / np.diag(A))
    n = A.shape[0]
    S = np.eye(n) - Minv @ A
    Ep = S @ (np.eye(n) - Pm @ np.linalg.solve(Ac, Pm.T @ A)) @ S
    B = (np.eye(n) - Ep) @ np.linalg.inv(A)
    b = np.random.default_rng(2).standard_normal(n)
    assert _rel(o.vcycle(b), B @ b) < 1e-10
    # the V(1,1) Jacobi cycle is a symmetric positive definite operator
    y = np.random.default_rng(3).standard_normal(n)
    assert abs(o.vcycle(b) @ y - b @ o.vcycle(y)) < 1e-10 * abs(b @ o.vcycle(b))
    assert b @ o.vcycle(b) > 0


def test_wcycle_closed_forms(oracle_mod):
    """W-cycle (Implementation/WCycle.m) pinned by what the recursion fixes:
    with two levels it is the V-cycle; from an initial guess x0 one cycle is
    the affine step x0 + B (b - A x0) of the zero-guess cycle B; three levels
    apply the level-1 two-level cycle twice, B_W = 2 B1 - B1 A1 B1 on the
    coarse residual; and the result stays symmetric positive definite."""
    prob = P.cantilever(8, 4, 4)
    x = P.random_density(8, 4, 4, 6)
    rng = np.random.default_rng(5)
    o2 = oracle_mod.Oracle(prob)
    o2.assemble(x)
    o2.setup_gmg(2, 0.6)
    b = rng.standard_normal(o2.nfree(0))
    assert np.array_equal(o2.wcycle(b), o2.vcycle(b))
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(3, 0.6)
    n0, n1 = o.nfree(0), o.nfree(1)
    b1 = rng.standard_normal(n1)
    x0 = rng.standard_normal(n1)
    # level 1 sits above the coarsest level: its W-cycle from x0 is the
    # two-level cycle's affine step
    lhs = o.wcycle(b1, x0=x0, level=1)
    rhs = x0 + o.vcycle(b1 - o.spmv(1, x0), level=1)
    assert _rel(lhs, rhs) < 1e-11
    # level 0: pre-smooth, then the coarse correction 2 B1 - B1 A1 B1
    A = o.csr("A", 0).toarray()
    Pm = o.csr("P", 0).toarray()
    A1 = o.csr("A", 1).toarray()
    B1 = np.column_stack([o.vcycle(e, level=1) for e in np.eye(n1)])
    BW = 2 * B1 - B1 @ A1 @ B1
    Minv = np.diag(0.6 / np.diag(A))
    S = np.eye(n0) - Minv @ A
    Ep = S @ (np.eye(n0) - Pm @ BW @ Pm.T @ A) @ S
    B = (np.eye(n0) - Ep) @ np.linalg.inv(A)
    b = rng.standard_normal(n0)
    w = o.wcycle(b)
    assert _rel(w, B @ b) < 1e-9
    assert _rel(w, o.vcycle(b)) > 1e-6            # three levels: W differs from V
    y = rng.standard_normal(n0)
    assert abs(w @ y - b @ o.wcycle(y)) < 1e-10 * abs(b @ w)
    assert b @ w > 0


def test_fmv_cycle_closed_forms(oracle_mod):
    """FMV cycle (Implementation/FMVCycle.m) pinned by its definition in
    closed form: the V-cycle from an initial guess x0 is the affine step
    x0 + B (b - A x0) of the zero-guess cycle; with two levels FMV is that
    step from x0 = P A_c^-1 P' b (exact coarse solve); with three levels the
    guess is P times the level-1 FMV of P' b, itself the level-1 step from
    P1 A_2^-1 P1' (P' b). The FMV solve reaches the direct solution."""
    prob = P.cantilever(8, 4, 4)
    x = P.random_density(8, 4, 4, 8)
    rng = np.random.default_rng(9)
    for L in (2, 3):
        o = oracle_mod.Oracle(prob)
        o.assemble(x)
        o.setup_gmg(L, 0.6)
        n0 = o.nfree(0)
        b = rng.standard_normal(n0)
        x0 = rng.standard_normal(n0)
        assert _rel(o.vcycle_x0(b, x0), x0 + o.vcycle(b - o.spmv(0, x0))) < 1e-11
        assert np.array_equal(o.vcycle_x0(b, None), o.vcycle(b))
        P0 = o.csr("P", 0).toarray()
        if L == 2:
            Ac = o.csr("A", 1).toarray()
            guess = P0 @ np.linalg.solve(Ac, P0.T @ b)
        else:
            P1 = o.csr("P", 1).toarray()
            A2 = o.csr("A", 2).toarray()
            b1 = P0.T @ b
            g1 = P1 @ np.linalg.solve(A2, P1.T @ b1)
            x1 = g1 + o.vcycle(b1 - o.spmv(1, g1), level=1)
            guess = P0 @ x1
        ref = guess + o.vcycle(b - o.spmv(0, guess))
        assert _rel(o.fmvcycle(b), ref) < 1e-9
        assert _rel(o.fmvcycle(b), o.vcycle(b)) > 1e-6
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 7)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(3, 0.6)
    u_d = sla.spsolve(o.csr("A", 0).tocsc(), prob.f)
    u_f, it_f, rr = o.solve(prob.f, cycle=2)
    assert rr <= 1e-8 and _rel(u_f, u_d) < 1e-6


def test_wcycle_solve_matches_direct(oracle_mod):
    prob = P.cantilever(16, 8, 8)
    x = P.random_density(16, 8, 8, 7)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(3, 0.6)
    u_v, it_v, _ = o.solve(prob.f)
    u_w, it_w, rr = o.solve(prob.f, cycle=1)
    assert rr <= 1e-8
    assert _rel(u_w, u_v) < 1e-6
    assert it_w <= it_v


def test_smooth_is_damped_jacobi(oracle_mod):
    prob = P.cantilever(4, 2, 2)
    o = oracle_mod.Oracle(prob)
    o.assemble(P.random_density(4, 2, 2, 5))
    o.setup_gmg(2, 0.6)
    A = o.csr("A", 0)
    M = o.jacobi_diag(0)
    assert _rel(M, A.diagonal() / 0.6) < 1e-15
    rng = np.random.default_rng(4)
    b, x0 = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    x = x0.copy()
    for _ in range(3):
        x = x + (b - A @ x) / M
    assert _rel(o.smooth(0, b, x0, 3), x) < 1e-14


@pytest.mark.parametrize("grid,L,seed", [((8, 4, 4), 3, 1), ((16, 8, 8), 3, 2), ((12, 4, 8), 2, 3)])
def test_solve_mg_matches_direct_solve(oracle_mod, grid, L, seed):
    prob = P.cantilever(*grid)
    x = P.random_density(*grid, seed)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(L, 0.6)
    u, it, rr = o.solve(prob.f)
    K = o.csr("A", 0).tocsc()
    ud = sla.spsolve(K, prob.f)
    assert rr <= 1e-8 and it > 1
    assert np.linalg.norm(prob.f - K @ u) <= 1e-8 * np.linalg.norm(prob.f) * 1.01
    assert _rel(u, ud) < 1e-6
    assert abs(prob.f @ u - prob.f @ ud) < 1e-9 * abs(prob.f @ ud)


def test_uniform_density_closed_form(oracle_mod):
    prob = P.cantilever(8, 4, 4)
    sols = {}
    for rho in (1.0, 0.3):
        o = oracle_mod.Oracle(prob)
        o.assemble(np.full(prob.nelem, rho))
        o.setup_gmg(3, 0.6)
        sols[rho] = o.solve(prob.f, tol=1e-12)[0]
    assert _rel(sols[0.3], sols[1.0] / (1e-9 + (1 - 1e-9) * 0.3 ** 3)) < 1e-9


def test_small_cantilever_fixture(oracle_mod):
    g = np.load(os.path.join(HERE, "golden", "small_cantilever.npz"))
    prob = P.cantilever(8, 4, 4)
    np.testing.assert_array_equal(prob.freedofs, g["freedofs"])
    o = oracle_mod.Oracle(prob)
    o.assemble(g["x"])
    o.setup_gmg(3, 0.6)
    u, it, _ = o.solve(prob.f)
    assert it == int(g["it"])
    assert _rel(u, g["u"]) < 1e-12
    K = o.csr("A", 0).tocsc()
    assert _rel(sla.spsolve(K, prob.f), g["u"]) < 1e-7


# ---------------------------------------------------------------------------
# gradient and filter
# ---------------------------------------------------------------------------
def test_grad_finite_differences_and_euler_identity(oracle_mod):
    prob = P.cantilever(4, 2, 2)
    x = P.random_density(4, 2, 2, 6)
    Emin, p = 1e-3, 3.0

    def compliance(xx):
        o = oracle_mod.Oracle(prob, Emin=Emin)
        o.assemble(xx)
        K = o.csr("A", 0).tocsc()
        u = sla.spsolve(K, prob.f)
        return prob.f @ u, u, o
    C0, u, o = compliance(x)
    g0 = o.grad0(prob.full(u), x)
    h = 1e-6
    for e in range(prob.nelem):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (compliance(xp)[0] - compliance(xm)[0]) / (2 * h)
        assert abs(fd - g0[e]) < 1e-5 * abs(g0).max(), e
    # Euler: sum_e x_e g0_e = -p (C - Emin sum_e u_e' Ke u_e)
    o1 = oracle_mod.Oracle(prob, Emin=0.0, p=1.0)
    ueKu = -o1.grad0(prob.full(u), np.ones(prob.nelem))  # u_e' Ke u_e
    assert abs(x @ g0 + p * (C0 - Emin * ueKu.sum())) < 1e-10 * abs(C0)


def test_filter_closed_form_weights(oracle_mod):
    W, w = oracle_mod.filter_matrix(5, 5, 5, 1.5, 2)
    e = 2 * 25 + 2 * 5 + 2  # interior element (2,2,2)
    row = W.getrow(e)
    vals = np.sort(row.data)
    assert row.nnz == 19
    edge = (1.5 - math.sqrt(2)) / 1.5
    np.testing.assert_allclose(vals, np.sort([1.0] + [1 / 3] * 6 + [edge] * 12), rtol=1e-15)
    assert abs(w[e] - (1 + 2 + 12 * edge)) < 1e-14
    assert abs(W - W.T).max() == 0
    np.testing.assert_allclose(w, np.asarray(W.sum(axis=1)).ravel(), rtol=1e-15)
    x = np.random.default_rng(0).random(125)
    y = np.random.default_rng(1).random(125)
    Wx = (W @ x) / w
    assert abs(Wx @ y - x @ (W.T @ (y / w))) < 1e-12
    assert np.allclose((W @ np.ones(125)) / w, 1.0)
    # opfilter 1: Gaussian weights exp(-d^2 alpha) beta, self weight beta
    Wg, _ = oracle_mod.filter_matrix(5, 5, 5, 1.5, 1)
    alpha, beta = 1 / (2 * (0.5 ** 2)), 1 / (2 * math.pi * 0.5)
    assert abs(Wg[e, e] - beta) < 1e-15
    assert abs(Wg[e, e + 1] - math.exp(-alpha) * beta) < 1e-15
    # the structured numpy filter agrees with the oracle's W
    assert _rel(P.apply_filter(x, 5, 5, 5, 1.5, 2), Wx) < 1e-14
    assert _rel(P.apply_filter(x, 5, 5, 5, 1.5, 2, trans=True), W.T @ (x / w)) < 1e-14


def test_filter_fixed_elements_decoupled(oracle_mod):
    fixed = np.zeros(64, dtype=np.uint8)
    fixed[[5, 21, 40]] = 1
    W, w = oracle_mod.filter_matrix(4, 4, 4, 1.5, 2, fixed=fixed)
    for e in (5, 21, 40):
        r = W.getrow(e)
        assert r.nnz == 1 and r[0, e] == 1.0
        assert W.getcol(e).nnz == 1


def test_problem_file_round_trip(tmp_path):
    """problem.from_mat reads a str struct (as the reference's problem files
    hold it) and rebuilds freedofs/f from its rows; a file whose stored
    vectors disagree with its rows is rejected."""
    import scipy.io as sio
    prob = P.from_rows(12, 4, 4, [[0, 0, 0, 4, 0, 4, 1, 1, 1]], [[12, 12, 0, 0, 0, 4, 0, -1, 0]])
    st = {"nelx": 12, "nely": 4, "nelz": 4, "l": 12.0, "h": 4.0, "w": 4.0, "E": 1.0, "nu": 0.3,
          "supp": np.array([[0, 0, 0, 4, 0, 4, 1, 1, 1]], float),
          "loads": np.array([[12, 12, 0, 0, 0, 4, 0, -1, 0]], float),
          "freedofs": prob.freedofs.astype(float), "f": prob.f}
    path = tmp_path / "p.mat"
    sio.savemat(path, {"str": st})
    q = P.from_mat(path)
    assert np.array_equal(q.freedofs, prob.freedofs) and np.array_equal(q.f, prob.f)
    st["f"] = prob.f * 2
    sio.savemat(path, {"str": st})
    with pytest.raises(ValueError):
        P.from_mat(path)
    assert P.from_mat(path, check=False).nfree == prob.nfree


@pytest.mark.skipif(not os.path.isdir("/root/reference/Implementation/Tests"), reason="reference problem files absent")
def test_reference_problem_files_read_back():
    """All 73 reference problem files load: those with rows rebuild their
    stored freedofs, f and fixed densities (checked inside from_mat), those
    without give their stored vectors, equal to the committed records."""
    stored = json.load(open(os.path.join(HERE, "golden", "stored_problems.json")))
    rows = json.load(open(os.path.join(HERE, "golden", "setup_fixtures.json")))
    assert len(stored) + len(rows) == 73
    for fx in stored + rows:
        p = P.from_mat(os.path.join("/root/reference/Implementation/Tests", fx["file"]))
        assert p.nfree == fx["nfree"] and _sha(p.freedofs) == fx["freedofs_sha256"]
        assert _sha(p.fixed_dens) == fx["fixed_dens_sha256"]
