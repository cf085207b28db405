// tmg_coarse_nd.inc.hpp -- coarsest-grid solve by nested dissection and a
// multifrontal Cholesky factorization (included by tmg_solver.hip inside its
// anonymous namespace, after the handle and the allocator).
//
// Implementation/SetupGMG.m:51-60 orders the coarsest Galerkin operator with
// amd and factors it with chol; Implementation/VCycle.m:35-36 solves with the
// two triangular factors. Here the ordering is a geometric nested dissection
// of the coarse node grid: the region is cut through its longest axis by a
// plane of nodes (the separator, which decouples the two halves of a 27-point
// stencil), recursively, down to leaves of at most 256 nodes. Each tree
// node is a front: its own free DOFs are the pivots P (elimination order:
// children before parents, nodes in grid order within a front), the free
// DOFs of the ancestor separators next to its region are its boundary B.
// The numeric factorization runs height by height on the device: assembly
// of the operator's entries and of the children's update matrices, then per
// front rocSOLVER/rocBLAS potrf, trsm, syrk (the update matrix), trtri and
// trmm, which leave M = [L_PP^-1 ; L_BP L_PP^-1] in the front's first p
// columns. All of it runs in order on the handle's stream: with the fronts
// of one height spread over concurrent side streams (each with its own
// rocBLAS handle), identical set-ups gave different factors from run to run,
// some of them not positive definite.
// The solve is two sweeps of matrix-vector products over the fronts of one
// height per launch (k_nd_fwd, k_nd_bwd).

int check_launch(tmg_solver *h);

struct NdHost {
    std::vector<NdFront> fr;
    std::vector<int64_t> P, B, src, dst, bsrc, tsrc;
    std::vector<int> cmap, hstart, ea_start, fw_start, bw_start;
    std::vector<int2> ea, fw, bw, tp;
    int64_t Fsize = 0, Tsize = 0, Mtsize = 0;
    int maxld = 0, nheight = 0;
};

struct NdNode {
    int box[3][2];     // region [lo, hi) per axis (x, y, z)
    int piv[3][2];     // own nodes
    int child[2] = {-1, -1};
    int height = 0;
};

// geometric nested dissection of the region; returns the tree node index
int nd_split(std::vector<NdNode> &t, const int box[3][2], int leaf) {
    const int idx = (int)t.size();
    t.emplace_back();
    int dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = box[a][1] - box[a][0];
    const int64_t nn = (int64_t)dims[0] * dims[1] * dims[2];
    int a = 0;
    for (int q = 1; q < 3; ++q)
        if (dims[q] > dims[a]) a = q;
    for (int q = 0; q < 3; ++q)
        for (int e = 0; e < 2; ++e) t[idx].box[q][e] = t[idx].piv[q][e] = box[q][e];
    if (nn <= leaf || dims[a] < 3) return idx;
    const int m = (box[a][0] + box[a][1]) / 2;
    int b1[3][2], b2[3][2];
    for (int q = 0; q < 3; ++q)
        for (int e = 0; e < 2; ++e) b1[q][e] = b2[q][e] = box[q][e];
    b1[a][1] = m;
    b2[a][0] = m + 1;
    t[idx].piv[a][0] = m;
    t[idx].piv[a][1] = m + 1;
    const int c0 = nd_split(t, b1, leaf);
    const int c1 = nd_split(t, b2, leaf);
    t[idx].child[0] = c0;
    t[idx].child[1] = c1;
    t[idx].height = 1 + std::max(t[c0].height, t[c1].height);
    return idx;
}

void nd_post(const std::vector<NdNode> &t, int i, std::vector<int> &order) {
    for (int c = 0; c < 2; ++c)
        if (t[i].child[c] >= 0) nd_post(t, t[i].child[c], order);
    order.push_back(i);
}

// Symbolic analysis of the coarsest level (structure only: it holds for every
// density). Returns false when the fronts do not fit the solve kernels' LDS.
bool nd_symbolic(const DevLevel &lc, int leaf, NdHost &o) {
    const int nx = lc.nx, ny = lc.ny, nz = lc.nz;
    const int64_t nn = (int64_t)nx * ny * nz, nc = lc.nfree;
    std::vector<int64_t> cidx(3 * nn, -1);
    for (int64_t q = 0; q < nc; ++q) cidx[lc.free_full[q]] = q;
    std::vector<NdNode> t;
    const int root_box[3][2] = {{0, nx}, {0, ny}, {0, nz}};
    nd_split(t, root_box, leaf);
    std::vector<int> post;
    nd_post(t, 0, post);
    // pivot DOFs and elimination ranks
    auto node_of = [&](int ix, int iy, int iz) { return (int64_t)iz * nx * ny + (int64_t)ix * ny + iy; };
    std::vector<std::vector<int64_t>> piv(t.size()), bnd(t.size());
    std::vector<int64_t> rank(nc, -1);
    int64_t r = 0;
    for (int i : post) {
        const NdNode &n = t[i];
        for (int iz = n.piv[2][0]; iz < n.piv[2][1]; ++iz)
            for (int ix = n.piv[0][0]; ix < n.piv[0][1]; ++ix)
                for (int iy = n.piv[1][0]; iy < n.piv[1][1]; ++iy)
                    for (int c = 0; c < 3; ++c) {
                        const int64_t q = cidx[3 * node_of(ix, iy, iz) + c];
                        if (q >= 0) { piv[i].push_back(q); rank[q] = r++; }
                    }
    }
    if (r != nc) return false;
    // boundary: free DOFs of the nodes next to the region (all on ancestor
    // separators), in elimination order
    for (int i : post) {
        const NdNode &n = t[i];
        std::vector<int64_t> &B = bnd[i];
        for (int iz = std::max(n.box[2][0] - 1, 0); iz < std::min(n.box[2][1] + 1, nz); ++iz)
            for (int ix = std::max(n.box[0][0] - 1, 0); ix < std::min(n.box[0][1] + 1, nx); ++ix)
                for (int iy = std::max(n.box[1][0] - 1, 0); iy < std::min(n.box[1][1] + 1, ny); ++iy) {
                    if (ix >= n.box[0][0] && ix < n.box[0][1] && iy >= n.box[1][0] && iy < n.box[1][1] &&
                        iz >= n.box[2][0] && iz < n.box[2][1])
                        continue;
                    for (int c = 0; c < 3; ++c) {
                        const int64_t q = cidx[3 * node_of(ix, iy, iz) + c];
                        if (q >= 0) B.push_back(q);
                    }
                }
        std::sort(B.begin(), B.end(), [&](int64_t a, int64_t b) { return rank[a] < rank[b]; });
    }
    // fronts in height order (postorder within a height)
    int H = 0;
    for (const NdNode &n : t) H = std::max(H, n.height);
    o.nheight = H + 1;
    std::vector<int> fid(t.size(), -1), byh;
    o.hstart.assign(o.nheight + 1, 0);
    for (int h = 0; h <= H; ++h) {
        o.hstart[h] = (int)byh.size();
        for (int i : post)
            if (t[i].height == h) { fid[i] = (int)byh.size(); byh.push_back(i); }
    }
    o.hstart[o.nheight] = (int)byh.size();
    const int nf = (int)byh.size();
    o.fr.resize(nf);
    int64_t foff = 0, toff = 0, poff = 0, boff = 0, coff = 0, moff = 0;
    for (int f = 0; f < nf; ++f) {
        const int i = byh[f];
        NdFront &F = o.fr[f];
        F.p = (int)piv[i].size();
        F.b = (int)bnd[i].size();
        F.ld = F.p + F.b;
        F.P_off = poff; F.B_off = boff; F.F_off = foff; F.T_off = toff; F.Mt_off = moff;
        poff += F.p; boff += F.b; foff += (int64_t)F.ld * F.ld; toff += F.b; moff += (int64_t)F.p * F.ld;
        for (int c = 0; c < 2; ++c) {
            F.child[c] = t[i].child[c] >= 0 ? fid[t[i].child[c]] : -1;
            F.cmap_off[c] = -1;
            if (F.child[c] >= 0) { F.cmap_off[c] = coff; coff += F.ld; }
        }
        o.P.insert(o.P.end(), piv[i].begin(), piv[i].end());
        o.B.insert(o.B.end(), bnd[i].begin(), bnd[i].end());
        o.maxld = std::max(o.maxld, F.ld);
    }
    o.Fsize = foff;
    o.Tsize = toff;
    o.Mtsize = moff;
    // child maps and operator entries
    o.cmap.assign(coff, -1);
    o.tsrc.assign(coff, -1);
    o.bsrc.resize(o.P.size());
    for (size_t k = 0; k < o.P.size(); ++k) o.bsrc[k] = lc.free_full[o.P[k]];
    std::vector<int> pos(nc, -1), cpos(nc, -1);
    for (int f = 0; f < nf; ++f) {
        const NdFront &F = o.fr[f];
        const int64_t *P = o.P.data() + F.P_off, *B = o.B.data() + F.B_off;
        auto dof = [&](int k) { return k < F.p ? P[k] : B[k - F.p]; };
        for (int k = 0; k < F.ld; ++k) pos[dof(k)] = k;
        for (int c = 0; c < 2; ++c) {
            if (F.child[c] < 0) continue;
            const NdFront &C = o.fr[F.child[c]];
            const int64_t *CB = o.B.data() + C.B_off;
            for (int k = 0; k < C.b; ++k) {
                if (pos[CB[k]] < 0) return false;    // a child's boundary lies in its parent's front
                cpos[CB[k]] = k;
            }
            for (int k = 0; k < F.ld; ++k) {
                const int ck = cpos[dof(k)];
                o.cmap[F.cmap_off[c] + k] = ck;
                o.tsrc[F.cmap_off[c] + k] = ck >= 0 ? C.T_off + ck : -1;
            }
            for (int k = 0; k < C.b; ++k) cpos[CB[k]] = -1;
        }
        // lower-triangle entries A(i, j), j a pivot, i in the front
        for (int pj = 0; pj < F.p; ++pj) {
            const int64_t gj = lc.free_full[P[pj]], nj = gj / 3;
            const int cj = (int)(gj % 3);
            const int jz = (int)(nj / ((int64_t)nx * ny)), jx = (int)((nj / ny) % nx), jy = (int)(nj % ny);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int iz = jz + dz, ix = jx + dx, iy = jy + dy;
                        if (iz < 0 || iz >= nz || ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
                        const int64_t ni = node_of(ix, iy, iz);
                        // stencil offset from node i to node j
                        const int k = (1 - dz) * 9 + (1 - dx) * 3 + (1 - dy);
                        for (int ci = 0; ci < 3; ++ci) {
                            const int64_t qi = cidx[3 * ni + ci];
                            if (qi < 0) continue;
                            const int pi = pos[qi];
                            if (pi < pj) continue;           // eliminated below, or the upper triangle
                            const int64_t src = k >= 13 ? sidx(ni, (k - 13) * 9 + ci * 3 + cj)
                                                        : sidx(nj, (13 - k) * 9 + cj * 3 + ci);
                            o.src.push_back(src);
                            o.dst.push_back(F.F_off + pi + (int64_t)pj * F.ld);
                        }
                    }
        }
        for (int k = 0; k < F.ld; ++k) pos[dof(k)] = -1;
    }
    // work lists per height
    o.ea_start.assign(o.nheight + 1, 0);
    o.fw_start.assign(o.nheight + 1, 0);
    o.bw_start.assign(o.nheight + 1, 0);
    for (int h = 0; h < o.nheight; ++h) {
        o.ea_start[h] = (int)o.ea.size();
        o.fw_start[h] = (int)o.fw.size();
        o.bw_start[h] = (int)o.bw.size();
        for (int f = o.hstart[h]; f < o.hstart[h + 1]; ++f) {
            const NdFront &F = o.fr[f];
            if (F.child[0] >= 0 || F.child[1] >= 0)
                for (int l = 0; l < F.ld; ++l) o.ea.push_back(make_int2(f, l));
            for (int r0 = 0; r0 < F.ld; r0 += 4) o.fw.push_back(make_int2(f, r0));
            for (int j = 0; j < F.p; ++j) o.tp.push_back(make_int2(f, j));
            for (int j0 = 0; j0 < F.p; j0 += 4) o.bw.push_back(make_int2(f, j0));
        }
    }
    o.ea_start[o.nheight] = (int)o.ea.size();
    o.fw_start[o.nheight] = (int)o.fw.size();
    o.bw_start[o.nheight] = (int)o.bw.size();
    // the solve kernels keep one front's right-hand side in LDS
    return (int64_t)o.maxld * 8 <= 48 * 1024;
}

// device copies of the symbolic structure (called once by the set-up)
int nd_upload(tmg_solver *h, NdHost &o) {
    h->nd_nfront = (int)o.fr.size();
    h->nd_maxld = o.maxld;
    h->nd_nheight = o.nheight;
    h->nd_fr = o.fr;
    h->nd_hstart = o.hstart;
    h->nd_ea_start = o.ea_start;
    h->nd_fw_start = o.fw_start;
    h->nd_bw_start = o.bw_start;
    h->nd_nent = (int64_t)o.src.size();
    h->nd_Fsize = o.Fsize;
    h->nd_ntp = (int64_t)o.tp.size();
    TRY(dalloc(h, &h->nd_fronts, (int64_t)o.fr.size()));
    TRY(dalloc(h, &h->nd_P, (int64_t)o.P.size()));
    TRY(dalloc(h, &h->nd_B, (int64_t)o.B.size()));
    TRY(dalloc(h, &h->nd_src, h->nd_nent));
    TRY(dalloc(h, &h->nd_dst, h->nd_nent));
    TRY(dalloc(h, &h->nd_cmap, (int64_t)o.cmap.size()));
    TRY(dalloc(h, &h->nd_tsrc, (int64_t)o.tsrc.size()));
    TRY(dalloc(h, &h->nd_bsrc, (int64_t)o.bsrc.size()));
    TRY(dalloc(h, &h->nd_info, (int64_t)o.fr.size()));
    TRY(dalloc(h, &h->nd_ea, (int64_t)o.ea.size()));
    TRY(dalloc(h, &h->nd_fw, (int64_t)o.fw.size()));
    TRY(dalloc(h, &h->nd_bw, (int64_t)o.bw.size()));
    TRY(dalloc(h, &h->nd_F, o.Fsize));
    TRY(dalloc(h, &h->nd_Mt, o.Mtsize));
    TRY(dalloc(h, &h->nd_tp, h->nd_ntp));
    TRY(dalloc(h, &h->nd_T, o.Tsize));
    TRY(dalloc(h, &h->nd_y, h->nc));
    TRY(dalloc(h, &h->nd_xc, h->nc));
    auto up = [&](void *d, const void *s, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, h->st) == hipSuccess;
    };
    if (!up(h->nd_fronts, o.fr.data(), sizeof(NdFront) * o.fr.size()) ||
        !up(h->nd_P, o.P.data(), sizeof(int64_t) * o.P.size()) || !up(h->nd_B, o.B.data(), sizeof(int64_t) * o.B.size()) ||
        !up(h->nd_src, o.src.data(), sizeof(int64_t) * o.src.size()) ||
        !up(h->nd_dst, o.dst.data(), sizeof(int64_t) * o.dst.size()) ||
        !up(h->nd_cmap, o.cmap.data(), sizeof(int) * o.cmap.size()) ||
        !up(h->nd_tsrc, o.tsrc.data(), sizeof(int64_t) * o.tsrc.size()) ||
        !up(h->nd_bsrc, o.bsrc.data(), sizeof(int64_t) * o.bsrc.size()) ||
        !up(h->nd_ea, o.ea.data(), sizeof(int2) * o.ea.size()) || !up(h->nd_fw, o.fw.data(), sizeof(int2) * o.fw.size()) ||
        !up(h->nd_bw, o.bw.data(), sizeof(int2) * o.bw.size()) || !up(h->nd_tp, o.tp.data(), sizeof(int2) * o.tp.size()) ||
        hipStreamSynchronize(h->st) != hipSuccess) {
        h->err = "upload of the coarse nested-dissection structure failed";
        return TMG_ERR_HIP;
    }
    return 0;
}

void nd_free(tmg_solver *h) {
    dfree(h, h->nd_xc); dfree(h, h->nd_y); dfree(h, h->nd_T); dfree(h, h->nd_tp); dfree(h, h->nd_Mt); dfree(h, h->nd_F);
    dfree(h, h->nd_bw);
    dfree(h, h->nd_fw); dfree(h, h->nd_ea); dfree(h, h->nd_info); dfree(h, h->nd_bsrc); dfree(h, h->nd_tsrc);
    dfree(h, h->nd_cmap); dfree(h, h->nd_dst);
    dfree(h, h->nd_src); dfree(h, h->nd_B); dfree(h, h->nd_P); dfree(h, h->nd_fronts);
}

// Numeric factorization of the coarsest operator (from its FP64 stencil).
int nd_factor(tmg_solver *h) {
    const DevLevel &lc = h->lev[h->L - 1];
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(h->nd_Fsize)), dim3(TPB), 0, h->st, h->nd_Fsize, 0.0, h->nd_F);
    hipLaunchKernelGGL((k_fill<int>), dim3(nblk(h->nd_nfront)), dim3(TPB), 0, h->st, (int64_t)h->nd_nfront, 0, h->nd_info);
    if (h->nd_nent > 0)
        hipLaunchKernelGGL(k_nd_assemble, dim3(nblk(h->nd_nent)), dim3(TPB), 0, h->st, h->nd_nent,
                           (const int64_t *)h->nd_src, (const int64_t *)h->nd_dst, (const double *)lc.S64, h->nd_F);
    TRY(check_launch(h));
    rocblas_set_stream(h->rb, h->st);
    const double one = 1.0, mone = -1.0;
    for (int hh = 0; hh < h->nd_nheight; ++hh) {
        const int e0 = h->nd_ea_start[hh], e1 = h->nd_ea_start[hh + 1];
        if (e1 > e0)
            hipLaunchKernelGGL(k_nd_extend_add, dim3((unsigned)(e1 - e0)), dim3(TPB), 0, h->st,
                               (const int2 *)h->nd_ea + e0, (const NdFront *)h->nd_fronts, (const int *)h->nd_cmap,
                               h->nd_F);
        TRY(check_launch(h));
        for (int f = h->nd_hstart[hh]; f < h->nd_hstart[hh + 1]; ++f) {
            const NdFront &F = h->nd_fr[f];
            if (F.p == 0) continue;
            rocblas_handle rb = h->rb;
            double *Fp = h->nd_F + F.F_off;
            double *FB = Fp + F.p;                              // rows B, columns P
            double *FBB = Fp + F.p + (int64_t)F.p * F.ld;       // rows B, columns B
            rocblas_status st = rocsolver_dpotrf(rb, rocblas_fill_lower, F.p, Fp, F.ld, h->nd_info + f);
            if (st == rocblas_status_success && F.b > 0) {
                st = rocblas_dtrsm(rb, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                   rocblas_diagonal_non_unit, F.b, F.p, &one, Fp, F.ld, FB, F.ld);
                if (st == rocblas_status_success)
                    st = rocblas_dsyrk(rb, rocblas_fill_lower, rocblas_operation_none, F.b, F.p, &mone, FB, F.ld, &one,
                                       FBB, F.ld);
            }
            if (st != rocblas_status_success) {
                h->err = "coarse multifrontal factorization: rocBLAS/rocSOLVER call failed (status " +
                         std::to_string((int)st) + ")";
                return TMG_ERR_HIP;
            }
        }
    }
    // pivot failures before the inverses of the diagonal factors
    std::vector<int> info(h->nd_nfront, 0);
    HIPX(hipMemcpyAsync(info.data(), h->nd_info, sizeof(int) * info.size(), hipMemcpyDeviceToHost, h->st));
    HIPX(hipStreamSynchronize(h->st));
    for (int f = 0; f < h->nd_nfront; ++f)
        if (info[f] != 0) {
            h->err = "coarse Galerkin operator is not positive definite (front " + std::to_string(f) + " potrf info " +
                     std::to_string(info[f]) + ")";
            return TMG_ERR_NOT_SPD;
        }
    for (int f = 0; f < h->nd_nfront; ++f) {
        const NdFront &F = h->nd_fr[f];
        if (F.p == 0) continue;
        rocblas_handle rb = h->rb;
        double *Fp = h->nd_F + F.F_off;
        rocblas_status st = rocsolver_dtrtri(rb, rocblas_fill_lower, rocblas_diagonal_non_unit, F.p, Fp, F.ld,
                                             h->nd_info + f);
        if (st == rocblas_status_success && F.b > 0)
            st = rocblas_dtrmm(rb, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                               rocblas_diagonal_non_unit, F.b, F.p, &one, Fp, F.ld, Fp + F.p, F.ld, Fp + F.p, F.ld);
        if (st != rocblas_status_success) {
            h->err = "coarse multifrontal factorization: rocBLAS/rocSOLVER call failed (status " +
                     std::to_string((int)st) + ")";
            return TMG_ERR_HIP;
        }
    }
    if (h->nd_ntp > 0)
        hipLaunchKernelGGL(k_nd_transpose, dim3((unsigned)h->nd_ntp), dim3(TPB), 0, h->st, (const int2 *)h->nd_tp,
                           (const NdFront *)h->nd_fronts, (const double *)h->nd_F, h->nd_Mt);
    return check_launch(h);
}

// x = A^-1 b on the coarsest level (b, x: level vectors of the cycle's type)
template <typename T>
void nd_solve(tmg_solver *h, const T *b, T *x, const int *done) {
    const size_t sh = sizeof(double) * (size_t)h->nd_maxld;
    for (int hh = 0; hh < h->nd_nheight; ++hh) {
        const int w0 = h->nd_fw_start[hh], w1 = h->nd_fw_start[hh + 1];
        hipLaunchKernelGGL((k_nd_fwd<T>), dim3((unsigned)(w1 - w0)), dim3(TPB), sh, h->st, (const int2 *)h->nd_fw + w0,
                           (const NdFront *)h->nd_fronts, (const int64_t *)h->nd_P, (const int64_t *)h->nd_bsrc,
                           (const int64_t *)h->nd_tsrc, (const double *)h->nd_Mt, b, h->nd_y, h->nd_T, done);
    }
    for (int hh = h->nd_nheight - 1; hh >= 0; --hh) {
        const int w0 = h->nd_bw_start[hh], w1 = h->nd_bw_start[hh + 1];
        hipLaunchKernelGGL((k_nd_bwd<T>), dim3((unsigned)(w1 - w0)), dim3(TPB), sh, h->st, (const int2 *)h->nd_bw + w0,
                           (const NdFront *)h->nd_fronts, (const int64_t *)h->nd_P, (const int64_t *)h->nd_B,
                           (const double *)h->nd_F, (const int64_t *)h->nd_bsrc, (const double *)h->nd_y, h->nd_xc, x,
                           done);
    }
}
