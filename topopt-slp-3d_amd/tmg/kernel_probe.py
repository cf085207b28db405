#!/usr/bin/env python3
"""Launch one solver kernel many times (or a device copy of known size) for
profiling with rocprofv3, e.g.

    rocprofv3 --pmc FETCH_SIZE -- python3 topopt-slp-3d_amd/tmg/kernel_probe.py --kernel 1 --precision mixed
    rocprofv3 --pmc FETCH_SIZE -- python3 topopt-slp-3d_amd/tmg/kernel_probe.py --copy-bytes 1073741824

The copy mode runs an elementwise scale (read n floats, write n floats) whose
HBM traffic is known, to calibrate the FETCH_SIZE / WRITE_SIZE counters.

Kernel ids are those of tmg_time_kernel: 0 FP64 fine apply (q = K d),
1 fine Jacobi sweep, 2 level-2 stencil sweep, 3 V-cycle, 4/5 fine fused
residual / post-sweep, 6 fine fused direction update + apply. The workload is
the 192x96x96 cantilever (filtered random densities, seed 3, 5 levels)
unless --grid/--levels say otherwise.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", type=int, default=1)
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--grid", default="192x96x96")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--copy-bytes", type=int, default=0, help="time a device-to-device copy of this size instead")
    a = ap.parse_args()
    import torch
    torch.cuda.init()
    if a.copy_bytes:
        # a plain streaming kernel of known traffic: read n floats, write n floats
        n = a.copy_bytes // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(a.reps):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.copy_bytes, "reps": a.reps}))
        return
    import tmg
    from tmg import problem as P
    nelx, nely, nelz = (int(v) for v in a.grid.split("x"))
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, a.seed)
    s = tmg.Solver(prob, a.levels, precision=a.precision)
    s.update_density(x)
    ms = s.time_kernel(a.kernel, a.reps)
    print(json.dumps({"kernel": a.kernel, "precision": a.precision, "avg_ms": ms, "reps": a.reps,
                      "n_dof": prob.ndof, "n_el": prob.nelem}))


if __name__ == "__main__":
    main()
