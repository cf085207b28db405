"""CPU pins of the oracle's multiresolution branch (mr.n = mr.d = 2, no GPU).

The oracle restates Implementation/ElemStiffMatrixGQMR.m (the 8 sub-element
integrands), Implementation/GlobalStiffMatrix.m:19-62 (the multiresolution
assembly) and Implementation/Grad.m:24-49 (the multiresolution gradient).
Each is checked here against something the restatement does not share:

* the integrands against an identity of the mathematics: the parent's
  trilinear field restricted to sub-box s is the sub-box's own trilinear
  field of the values P_s u at its corners, and a Q1 element matrix scales
  with the box size, so k{s} = P_s' Ke(box / 2) P_s with Ke the closed form
  (itself pinned to Implementation/ElemStiffMatrix.m). A mirrored or permuted
  sub-element order fails it;
* the assembly against an independent scipy assembly that follows the
  reference's own density numbering, transcribed loop by loop from
  Implementation/StartMR.m:58-69 (strDens.fde) and Implementation/Grad.m:28-40
  (the sorted sub-element indices v paired with k{i2});
* uniform sub-densities against the single-resolution operator;
* the gradient against central finite differences of the compliance and the
  Euler identity of the SIMP interpolation.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from tmg import problem as P

BOXES = [(1.0, 1.0, 1.0), (1.0, 0.7, 1.3), (2.0, 1.0, 0.5)]


def _sub_prolongation(s):
    """24 x 24 trilinear map from the parent element's corner values to the
    corners of sub-box s = sy + 2 sx + 4 sz (local node a = dy + 2 dx + 4 dz)."""
    sy, sx, sz = s & 1, (s >> 1) & 1, s >> 2
    Pm = np.zeros((24, 24))
    for a in range(8):
        pos = np.array([sx + ((a >> 1) & 1), sy + (a & 1), sz + (a >> 2)]) / 2.0
        for b in range(8):
            corner = np.array([(b >> 1) & 1, b & 1, b >> 2])
            w = np.prod(1.0 - np.abs(pos - corner))
            for c in range(3):
                Pm[3 * a + c, 3 * b + c] = w
    return Pm


def _integrands_from_identity(oracle_mod, el, eh, ew, nu=0.3):
    Kh = oracle_mod.elem_stiff(el / 2, eh / 2, ew / 2, nu)
    return np.array([_sub_prolongation(s).T @ Kh @ _sub_prolongation(s) for s in range(8)])


@pytest.mark.parametrize("box", BOXES)
def test_mr_integrands_are_half_box_galerkin_of_ke(oracle_mod, box):
    k = oracle_mod.elem_stiff_mr(2, *box, 0.3)
    ref = _integrands_from_identity(oracle_mod, *box)
    scale = np.abs(ref).max()
    assert np.abs(k - ref).max() <= 1e-14 * scale
    # the check tells sub-element orders apart: a mirrored order (s <-> s ^ 7)
    # or an x/y swap in the index (s = sx + 2 sy + 4 sz) is far off
    mirrored = k[[s ^ 7 for s in range(8)]]
    assert np.abs(mirrored - ref).max() > 0.1 * scale
    swap = [((s >> 1) & 1) | ((s & 1) << 1) | (s & 4) for s in range(8)]
    if box[0] != box[1]:
        assert np.abs(k[swap] - ref).max() > 0.01 * scale
    # and the half-box closed form is half the full one (Ke(h el) = h Ke)
    assert np.abs(oracle_mod.elem_stiff(box[0] / 2, box[1] / 2, box[2] / 2, 0.3)
                  - 0.5 * oracle_mod.elem_stiff(*box, 0.3)).max() <= 1e-15 * scale * 10


def _fde_transcribed(nelx, nely, nelz, n):
    """strDens.fde, the 1-based first density element of each finite element,
    generated as Implementation/StartMR.m:58-69 does."""
    nelyn, nelxyn = nely * n, nelx * n * nely * n
    fde = np.zeros(nelx * nely * nelz, dtype=np.int64)
    jn = 1
    for i in range(nelx * nely * nelz):
        fde[i] = jn
        jn += n
        if (jn - 1) % nelyn == 0:
            jn += (n - 1) * nelyn
            if jn % nelxyn == 1:
                jn += (n - 1) * nelxyn
    return fde


def _density_indices(nelx, nely, nelz, n):
    """v of Implementation/Grad.m:28-40 for every element: the sorted 1-based
    density indices whose i2-th entry pairs with the integrand k{i2}."""
    fde = _fde_transcribed(nelx, nely, nelz, n)
    nely_d, nelxy_d = nely * n, nelx * n * nely * n
    out = np.zeros((fde.size, n ** 3), dtype=np.int64)
    for i, f in enumerate(fde):
        vb = [f + n1 * nelxy_d + n2 * nely_d for n1 in range(n) for n2 in range(n)]
        v = [b + n3 for n3 in range(n) for b in vb]
        out[i] = np.sort(v)
    return out


def _assemble_mr_scipy(oracle_mod, prob, xd, Emin=1e-9, p=3.0):
    """K(free, free) = sum_e sum_i2 (Emin + (1 - Emin) xd[v_i2]^p) k{i2}, with k
    from the identity above and v from the transcribed density map."""
    k = _integrands_from_identity(oracle_mod, prob.el, prob.eh, prob.ew, prob.nu)
    V = _density_indices(prob.nelx, prob.nely, prob.nelz, 2) - 1
    Ed = Emin + (1.0 - Emin) * xd ** p
    nx, ny = prob.nelx + 1, prob.nely + 1
    rows, cols, vals = [], [], []
    e = 0
    for ez in range(prob.nelz):
        for ex in range(prob.nelx):
            for ey in range(prob.nely):
                dofs = np.array([3 * ((ez + (a >> 2)) * nx * ny + (ex + ((a >> 1) & 1)) * ny + ey + (a & 1)) + c
                                 for a in range(8) for c in range(3)])
                Ke = np.einsum("s,sij->ij", Ed[V[e]], k)
                rows.append(np.repeat(dofs, 24))
                cols.append(np.tile(dofs, 24))
                vals.append(Ke.ravel())
                e += 1
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(prob.ndof, prob.ndof))
    fd = prob.freedofs - 1
    return K[fd][:, fd]


def test_density_map_transcription_closed_form():
    """The transcribed fde loop gives the closed form the kernels use:
    sub-element s = sy + 2 sx + 4 sz of element (ex, ey, ez) is density
    element (2 ez + sz, 2 ex + sx, 2 ey + sy) of the 2nelx x 2nely x 2nelz mesh."""
    nelx, nely, nelz = 3, 2, 2
    V = _density_indices(nelx, nely, nelz, 2) - 1
    NX, NY = 2 * nelx, 2 * nely
    e = 0
    for ez in range(nelz):
        for ex in range(nelx):
            for ey in range(nely):
                for s in range(8):
                    sy, sx, sz = s & 1, (s >> 1) & 1, s >> 2
                    assert V[e, s] == (2 * ez + sz) * NX * NY + (2 * ex + sx) * NY + 2 * ey + sy
                e += 1
    assert np.array_equal(np.sort(V.ravel()), np.arange(8 * nelx * nely * nelz))


@pytest.mark.parametrize("box", [(1.0, 1.0, 1.0), (1.0, 0.7, 1.3)])
def test_assemble_mr_matches_independent_assembly(oracle_mod, box):
    prob = P.cantilever(4, 2, 2)
    prob.el, prob.eh, prob.ew = box
    xd = np.random.default_rng(3).uniform(1e-3, 1.0, 8 * prob.nelem)
    o = oracle_mod.Oracle(prob)
    o.assemble_mr(2, xd)
    K = o.csr("A", 0)
    Kn = _assemble_mr_scipy(oracle_mod, prob, xd)
    assert abs(K - Kn).max() <= 1e-14 * abs(Kn).max()


def test_assemble_mr_uniform_subdensities_equal_single_resolution(oracle_mod):
    nelx, nely, nelz = 4, 2, 2
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, 5)
    V = _density_indices(nelx, nely, nelz, 2) - 1
    xd = np.zeros(8 * prob.nelem)
    for e in range(prob.nelem):
        xd[V[e]] = x[e]
    o = oracle_mod.Oracle(prob)
    o.assemble_mr(2, xd)
    Kmr = o.csr("A", 0)
    o.assemble(x)
    K = o.csr("A", 0)
    assert abs(Kmr - K).max() <= 1e-14 * abs(K).max()


def test_grad0_mr_finite_differences_and_euler_identity(oracle_mod):
    """Implementation/Grad.m:24-49 (mr.interp = false): the unfiltered
    gradient on the density mesh is dC/dxd by central differences of the
    compliance C = f'u (direct solves), and sum_v xd_v g_v equals
    -p (C - Emin sum_v u_e' k{i2} u_e)."""
    prob = P.cantilever(2, 1, 1)
    Emin, p = 1e-3, 3.0
    xd = np.random.default_rng(9).uniform(0.2, 1.0, 8 * prob.nelem)

    def compliance(xx):
        o = oracle_mod.Oracle(prob, Emin=Emin, p=p)
        o.assemble_mr(2, xx)
        u = sla.spsolve(o.csr("A", 0).tocsc(), prob.f)
        return float(prob.f @ u), u, o

    C0, u, o = compliance(xd)
    g = o.grad0_mr(2, prob.full(u), xd)
    # the scipy assembly of the same operator solves to the same compliance
    un = sla.spsolve(_assemble_mr_scipy(oracle_mod, prob, xd, Emin=Emin, p=p).tocsc(), prob.f)
    assert abs(prob.f @ un - C0) <= 1e-12 * abs(C0)
    h = 1e-6
    for v in range(xd.size):
        xp, xm = xd.copy(), xd.copy()
        xp[v] += h
        xm[v] -= h
        fd = (compliance(xp)[0] - compliance(xm)[0]) / (2 * h)
        assert abs(fd - g[v]) <= 1e-6 * abs(g).max(), v
    # Euler identity: with p = 1 and Emin = 0 the oracle's g is -u_e' k u_e
    o1 = oracle_mod.Oracle(prob, Emin=0.0, p=1.0)
    o1.assemble_mr(2, np.ones_like(xd))
    uku = -o1.grad0_mr(2, prob.full(u), np.ones_like(xd))
    k = _integrands_from_identity(oracle_mod, 1.0, 1.0, 1.0)
    V = _density_indices(2, 1, 1, 2) - 1
    uf = prob.full(u)
    nx, ny = 3, 2
    for e in range(prob.nelem):
        ex = e // 1
        dofs = np.array([3 * (((a >> 2)) * nx * ny + (ex + ((a >> 1) & 1)) * ny + (a & 1)) + c
                         for a in range(8) for c in range(3)])
        ue = uf[dofs]
        for s in range(8):
            assert abs(uku[V[e, s]] - ue @ k[s] @ ue) <= 1e-12 * abs(uku).max()
    assert abs(xd @ g + p * (C0 - Emin * uku.sum())) <= 1e-10 * abs(C0)
