"""CPU pins of the spectral element form the fine-level and level-1 kernels use
(no GPU): the compile-time nonzero patterns WHT_NZ / WHT_MR_NZ of the kernel
header against the oracle's element matrices, the transform identities against
direct products, and the level-1 element decomposition against the oracle's
assembled Galerkin operator P'AP (Implementation/SetupGMG.m:47-49)."""
import os
import re

import numpy as np
import pytest

from tmg import problem as P

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "topopt-slp-3d_amd", "csrc",
                   "tmg_kernels.hpp")
BIT = [2, 1, 4]          # e_c: bit of displacement component c (x, y, z) in the local node index a = dy + 2dx + 4dz


def _H():
    return np.array([[(-1) ** bin(k & b).count("1") for b in range(8)] for k in range(8)], float)


def _header_tables():
    txt = open(HDR).read()
    m = re.search(r"WHT_NZ\[3\]\[3\] = \{(.*?)\};", txt, re.S)
    nz = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", m.group(1))]
    m = re.search(r"WHT_MR_NZ\[24\] = \{(.*?)\};", txt, re.S)
    mr = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", m.group(1))]
    return np.array(nz).reshape(3, 3), mr


def _coupling(K):
    """C[c][d][k][l] = 1/64 sum_ab H[k][a] K[3a+c][3b+d] H[l][b]: y = H' C H x with
    the kernels' unscaled inverse transform."""
    H = _H()
    return np.einsum("ka,acbd,lb->cdkl", H, K.reshape(8, 3, 8, 3), H) / 64


def _spectral_apply(C, x):
    H = _H()
    xh = np.einsum("kb,bd->dk", H, x.reshape(8, 3))
    yh = np.einsum("cdkl,dl->ck", C, xh)
    return np.einsum("ka,ck->ac", H, yh).reshape(-1)


BOXES = [(1.0, 1.0, 1.0), (1.0, 0.7, 1.3), (2.0, 1.0, 0.5)]


@pytest.mark.parametrize("box", BOXES)
def test_fine_element_spectral_form(oracle_mod, box):
    """Ke of a box couples frequency k of component c only with k ^ e_c ^ e_d of
    component d, on exactly the 45 (c, d, k) of WHT_NZ."""
    Ke = oracle_mod.elem_stiff(*box, 0.3)
    C = _coupling(Ke)
    scale = np.abs(C).max()
    nz_hdr, _ = _header_tables()
    for c in range(3):
        for d in range(3):
            for k in range(8):
                for l in range(8):
                    on = l == (k ^ BIT[c] ^ BIT[d]) and (nz_hdr[c, d] >> k) & 1
                    assert (abs(C[c, d, k, l]) > 1e-12 * scale) == bool(on), (c, d, k, l)
    x = np.random.default_rng(1).standard_normal(24)
    y = _spectral_apply(C, x)
    assert np.abs(y - Ke @ x).max() <= 1e-14 * np.abs(Ke @ x).max()


def _mr_table(Ks):
    """A[c][d][k][l] with C_s = A H[m][s], m = k ^ l ^ e_c ^ e_d, and the largest
    deviation from that structure."""
    H = _H()
    Cs = np.array([_coupling(K) for K in Ks])
    A = np.zeros((3, 3, 8, 8))
    err = 0.0
    for c in range(3):
        for d in range(3):
            for k in range(8):
                for l in range(8):
                    m = k ^ l ^ BIT[c] ^ BIT[d]
                    A[c, d, k, l] = Cs[:, c, d, k, l] @ H[m] / 8
                    err = max(err, np.abs(Cs[:, c, d, k, l] - A[c, d, k, l] * H[m]).max())
    return A, err / np.abs(Cs).max()


def _check_mr_pattern(A):
    _, mr_hdr = _header_tables()
    scale = np.abs(A).max()
    for c in range(3):
        for k in range(8):
            for d in range(3):
                for l in range(8):
                    on = (mr_hdr[8 * c + k] >> (8 * d + l)) & 1
                    assert (abs(A[c, d, k, l]) > 1e-12 * scale) == bool(on), (c, k, d, l)
    assert sum(bin(v).count("1") for v in mr_hdr) == 279


def _galerkin_integrands(Ke):
    """P_s' Ke P_s for the 8 children s = sy + 2sx + 4sz of a coarse element,
    P_s the trilinear weights (Implementation/Prolongation.m:15-37) from the
    coarse element's nodes to child s's nodes."""
    out = np.zeros((8, 24, 24))
    for s in range(8):
        sy, sx, sz = s & 1, (s >> 1) & 1, s >> 2
        Pm = np.zeros((24, 24))
        for a in range(8):
            pos = np.array([sx + ((a >> 1) & 1), sy + (a & 1), sz + (a >> 2)]) / 2.0
            for b in range(8):
                w = np.prod(1 - np.abs(pos - np.array([(b >> 1) & 1, b & 1, b >> 2])))
                for c in range(3):
                    Pm[3 * a + c, 3 * b + c] = w
        out[s] = Pm.T @ Ke @ Pm
    return out


@pytest.mark.parametrize("box", BOXES)
@pytest.mark.parametrize("kind", ["multiresolution", "galerkin"])
def test_eight_integrand_spectral_form(oracle_mod, box, kind):
    """The 8 sub-element (multiresolution, Implementation/ElemStiffMatrixGQMR.m)
    or child Galerkin integrands: each coupling is one Hadamard mode of the 8
    moduli times a table A on the WHT_MR_NZ pattern, and the transform
    reproduces sum_s E_s K_s x."""
    Ke = oracle_mod.elem_stiff(*box, 0.3)
    if kind == "multiresolution":
        Ks = np.asarray(oracle_mod.elem_stiff_mr(2, *box, 0.3)).reshape(8, 24, 24)
        assert np.abs(Ks.sum(0) - Ke).max() <= 1e-13 * np.abs(Ke).max()
    else:
        Ks = _galerkin_integrands(Ke)
    A, err = _mr_table(Ks)
    assert err < 1e-13
    _check_mr_pattern(A)
    rng = np.random.default_rng(2)
    E = rng.uniform(1e-3, 1.0, 8)
    x = rng.standard_normal(24)
    H = _H()
    Eh = H @ E
    C = np.zeros((3, 3, 8, 8))
    for c in range(3):
        for d in range(3):
            for k in range(8):
                for l in range(8):
                    C[c, d, k, l] = A[c, d, k, l] * Eh[k ^ l ^ BIT[c] ^ BIT[d]]
    yref = np.einsum("s,sij->ij", E, Ks) @ x
    assert np.abs(_spectral_apply(C, x) - yref).max() <= 1e-13 * np.abs(yref).max()


def test_level1_element_form_matches_oracle_galerkin(oracle_mod):
    """With the supports on the plane x = 0 (even), every fixed fine DOF
    interpolates only from fixed coarse DOFs, and the reference's Galerkin
    operator P(free,free)' A(free,free) P(free,free) equals the sum over coarse
    elements of sum_s E_s P_s' Ke P_s with the coarse fixed DOFs removed."""
    nelx, nely, nelz = 4, 2, 2
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, 7)
    o = oracle_mod.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(2, 0.6)
    E = 1e-9 + (1 - 1e-9) * x ** 3
    Kt = _galerkin_integrands(oracle_mod.elem_stiff(1, 1, 1, 0.3))
    cx, cy, cz = nelx // 2 + 1, nely // 2 + 1, nelz // 2 + 1
    nc = 3 * cx * cy * cz
    A1 = np.zeros((nc, nc))
    for ez in range(nelz // 2):
        for ex in range(nelx // 2):
            for ey in range(nely // 2):
                Ke_c = np.zeros((24, 24))
                for s in range(8):
                    fx, fy, fz = 2 * ex + ((s >> 1) & 1), 2 * ey + (s & 1), 2 * ez + (s >> 2)
                    Ke_c += E[fz * nelx * nely + fx * nely + fy] * Kt[s]
                nodes = [(ez + (a >> 2)) * cx * cy + (ex + ((a >> 1) & 1)) * cy + ey + (a & 1) for a in range(8)]
                dofs = np.array([3 * n + c for n in nodes for c in range(3)])
                A1[np.ix_(dofs, dofs)] += Ke_c
    free1 = o.level_free(1)
    A1f = A1[np.ix_(free1, free1)]
    v = np.random.default_rng(8).standard_normal(free1.size)
    y = o.spmv(1, v)
    assert np.abs(A1f @ v - y).max() <= 1e-12 * np.abs(y).max()
