#!/usr/bin/env python3
"""Mint the golden fixtures under tests/golden/ (run in the build container,
where the reference tree is mounted; the GPU box never reads the reference).

1. setup_fixtures.json -- problem-setup known answers from the reference's
   own problem library (Implementation/Tests/*.mat, MATLAB v5 files read with
   scipy.io.loadmat, which executes nothing): for every file that keeps its
   support and load rows, the rows, the fixed-density rows (str.fix), and the
   resulting free-DOF list (count + SHA-256 of its 1-based int64 bytes), the
   non-zero entries of the load vector str.f and the fixed elements
   (str.fixedDens, count + SHA-256, solid count). stored_problems.json: the
   same record (without rows) for the files saved without rows.
2. small_cantilever.npz -- oracle outputs on an 8x4x4 cantilever (filtered
   random densities, seed 11, 3 levels): u, PCG count, compliance and
   gradient, each cross-checked here against an independent scipy solve.
3. oracle_<grid>_seed<s>.json -- FP64 oracle compliance / iterations of the
   larger benchmark workloads (with --large), used by bench.py to state each
   GPU variant's error against the oracle.
"""
import argparse
import glob
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "topopt-slp-3d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF_TESTS = "/root/reference/Implementation/Tests"


def fd_hash(fd1):
    return hashlib.sha256(np.ascontiguousarray(fd1, dtype="<i8").tobytes()).hexdigest()


def setup_fixtures():
    out, stored = [], []
    for path in sorted(glob.glob(os.path.join(REF_TESTS, "*.mat"))):
        d = sio_load(path)
        s = d["str"]
        names = set(s._fieldnames)
        fd = np.asarray(s.freedofs, dtype=np.int64).ravel()
        f = np.asarray(s.f, dtype=float).ravel()
        nz = np.nonzero(f)[0]
        fdens = np.asarray(getattr(s, "fixedDens", []), dtype=np.int64).ravel()
        fval = np.asarray(getattr(s, "fixedDensVal", []), dtype=float).ravel()
        rec = {"file": os.path.basename(path), "nelx": int(s.nelx), "nely": int(s.nely), "nelz": int(s.nelz),
               "l": float(s.l), "h": float(s.h), "w": float(s.w),
               "nfree": int(fd.size), "freedofs_sha256": fd_hash(fd),
               "fixed_dens_count": int(fdens.size), "fixed_dens_sha256": fd_hash(fdens),
               "fixed_dens_solid": int(np.sum(fval == 1))}
        if "fix" in names and np.asarray(s.fix).size:
            rec["fix"] = np.atleast_2d(np.asarray(s.fix, dtype=float)).tolist()
        if "supp" in names and "loads" in names:
            rec.update(supp=np.atleast_2d(np.asarray(s.supp, dtype=float)).tolist(),
                       loads=np.atleast_2d(np.asarray(s.loads, dtype=float)).tolist(),
                       f_nonzero_index1=(nz + 1).tolist(), f_nonzero_value=f[nz].tolist())
            out.append(rec)
        else:
            rec.update(f_nonzero_count=int(nz.size), f_sum=float(f.sum()))
            stored.append(rec)
    with open(os.path.join(HERE, "setup_fixtures.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    with open(os.path.join(HERE, "stored_problems.json"), "w") as fh:
        json.dump(stored, fh, indent=1)
    print(f"setup_fixtures.json: {len(out)} problems with rows; stored_problems.json: {len(stored)} without rows")


def sio_load(path):
    import scipy.io as sio
    return sio.loadmat(path, squeeze_me=True, struct_as_record=False)


def small_cantilever():
    import scipy.sparse.linalg as sla

    import oracle
    from tmg import problem as P
    prob = P.cantilever(8, 4, 4)
    x = P.random_density(8, 4, 4, 11)
    o = oracle.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(3, 0.6)
    u, it, rr = o.solve(prob.f)
    K = o.csr("A", 0).tocsc()
    ud = sla.spsolve(K, prob.f)
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-7
    W, w = oracle.filter_matrix(8, 4, 4, 1.5, 2)
    g0 = o.grad0(prob.full(u), x)
    g = W.T @ (g0 / w)
    np.savez(os.path.join(HERE, "small_cantilever.npz"), x=x, u=u, it=it, compliance=prob.f @ u, g0=g0, g=g,
             freedofs=prob.freedofs, f=prob.f)
    print(f"small_cantilever.npz: it={it} C={prob.f @ u:.12e}")


def large(grid, seed, levels):
    import time

    import oracle
    from tmg import problem as P
    nelx, nely, nelz = grid
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, seed) if seed is not None else np.full(prob.nelem, 0.3)
    t0 = time.time()
    o = oracle.Oracle(prob)
    o.assemble(x)
    o.setup_gmg(levels, 0.6)
    t1 = time.time()
    u, it, rr = o.solve(prob.f)
    t2 = time.time()
    r = prob.f - o.spmv(0, u)
    # a fixed sample of u: 256 evenly spaced free DOFs and every loaded DOF
    idx = np.unique(np.concatenate([np.linspace(0, prob.nfree - 1, 256).astype(np.int64),
                                    np.nonzero(prob.f)[0]]))
    rec = {"grid": list(grid), "seed": seed, "levels": levels, "nfree": int(prob.nfree), "it": int(it),
           "relres": float(rr), "true_relres": float(np.linalg.norm(r) / np.linalg.norm(prob.f)),
           "compliance": float(prob.f @ u), "u_norm": float(np.linalg.norm(u)),
           "u_tip_mean_y": float(np.mean(u[np.nonzero(prob.f)[0]])),
           "u_sample_index": idx.tolist(), "u_sample": u[idx].tolist(),
           "setup_s": t1 - t0, "solve_s": t2 - t1, "omp_threads": os.environ.get("OMP_NUM_THREADS")}
    name = f"oracle_{nelx}x{nely}x{nelz}_seed{seed}.json"
    with open(os.path.join(HERE, name), "w") as fh:
        json.dump(rec, fh, indent=1)
    print(name, rec)


def void_heavy(iters=30):
    """Void-heavy designs from the oracle's own SLP loop (volfrac 0.3, linear
    filter rmin 1.5): 48x24x24 single resolution (4 levels) and 24x12x12 with
    2x2x2 density sub-elements (3 levels). For each: the filtered design, the
    fraction of exactly-void density elements, the oracle's FP64 V-cycle and
    W-cycle iteration counts and compliance at tol 1e-8, and per-level
    operator probes (||A_l v||, v'A_l v for seeded v) for the GPU test."""
    import time

    import oracle
    import slp_oracle
    from tmg import problem as P
    out = {}
    arrays = {}
    for name, grid, mr, L in (("sr48", (48, 24, 24), 1, 4), ("mr24", (24, 12, 12), 2, 3)):
        prob = P.cantilever(*grid)
        t0 = time.time()
        hist, x, xfil, u = slp_oracle.str_slp(prob, 0.3, ngrids=L, mr=mr, maxiter=iters)
        t_slp = time.time() - t0
        o = oracle.Oracle(prob)
        if mr == 1:
            o.assemble(xfil)
        else:
            o.assemble_mr(mr, xfil)
        o.setup_gmg(L, 0.6)
        rec = {"grid": list(grid), "mr": mr, "levels": L, "slp_iterations": iters, "slp_s": t_slp,
               "void_fraction": float(np.mean(xfil == 0.0)), "min_positive": float(xfil[xfil > 0].min()),
               "slp_pcg_iterations": [int(h["pcgit"]) for h in hist]}
        for cyc, cname in ((0, "V"), (1, "W")):
            t0 = time.time()
            uu, it, rr = o.solve(prob.f, cycle=cyc, maxiter=5000)
            r = prob.f - o.spmv(0, uu)
            rec[cname] = {"it": int(it), "relres": float(rr),
                          "true_relres": float(np.linalg.norm(r) / np.linalg.norm(prob.f)),
                          "compliance": float(prob.f @ uu), "solve_s": time.time() - t0}
        probes = []
        for lev in range(L):
            v = np.random.default_rng(100 + lev).standard_normal(o.nfree(lev))
            Av = o.spmv(lev, v)
            probes.append({"norm_Av": float(np.linalg.norm(Av)), "vAv": float(v @ Av)})
        rec["level_probes"] = probes
        out[name] = rec
        arrays[name + "_xfil"] = xfil
        print(name, {k: v for k, v in rec.items() if k != "slp_pcg_iterations"})
    with open(os.path.join(HERE, "void_heavy.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "void_heavy.npz"), **arrays)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true", help="also the 96x48x48 and 192x96x96 oracle solves")
    ap.add_argument("--only", default=None, help="only this large grid (e.g. 192x96x96)")
    ap.add_argument("--void", action="store_true", help="only the void-heavy SLP designs (void_heavy.json/.npz)")
    a = ap.parse_args()
    if a.void:
        void_heavy()
        sys.exit(0)
    if a.only:
        g = tuple(int(v) for v in a.only.split("x"))
        large(g, {96: 2, 192: 3}[g[0]], {96: 4, 192: 5}[g[0]])
        sys.exit(0)
    setup_fixtures()
    small_cantilever()
    if a.large:
        large((96, 48, 48), 2, 4)
        large((192, 96, 96), 3, 5)
