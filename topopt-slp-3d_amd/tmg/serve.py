"""Continuous-batching service of equilibrium solves (libtmg's tmg_server_*).

``Server`` queues solve requests of uneven grid sizes; the native server
admits them as device memory and slots free up, overlaps their PCG
iterations on the device and reports per-request timings on its own clock.
``stream_workload`` builds the benchmark's request stream (cantilevers of
32x16x16 ... 128x64x64 elements, filtered random densities, random tip
loads, Poisson arrivals) and ``run_stream`` replays it in real time.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from . import _lib
from . import problem as P
from ._lib import TMG_MIXED, TmgError, lib
from .gmg import _PREC, _dp, _f64


def _structs(prob, ngrids, omega, nu1, nu2, tol, maxiter, p, emin):
    fd = np.ascontiguousarray(prob.freedofs, dtype=np.int64)
    pr = _lib.Problem_t(prob.nelx, prob.nely, prob.nelz, prob.el, prob.eh, prob.ew, float(prob.E), prob.nu,
                        float(p), float(emin), fd.ctypes.data_as(C.POINTER(C.c_int64)), fd.size)
    mg = _lib.MgParams_t(ngrids, 0, 0, omega, nu1, nu2, tol, maxiter, 0, 60000)
    return fd, pr, mg


def estimate_bytes(prob, ngrids, precision="mixed"):
    fd, pr, mg = _structs(prob, ngrids, 0.6, 1, 1, 1e-8, 5000, 3.0, 1e-9)
    return int(lib().tmg_server_estimate(C.byref(pr), C.byref(mg), _PREC[precision]))


class Server:
    def __init__(self, budget_bytes=200e9, max_active=8):
        h = C.c_void_p()
        rc = lib().tmg_server_create(float(budget_bytes), int(max_active), C.byref(h))
        if rc != 0:
            raise TmgError(rc, _lib.last_error(None))
        self.h = h
        self._meta = {}

    def _err(self):
        m = lib().tmg_server_last_error(self.h)
        return m.decode() if m else ""

    def submit(self, prob, x, ngrids, precision="mixed", f=None, x0=None, t_arrival=-1.0, omega=0.6,
               nu1=1, nu2=1, tol=1e-8, maxiter=5000, p=3.0, emin=1e-9):
        fd, pr, mg = _structs(prob, ngrids, omega, nu1, nu2, tol, maxiter, p, emin)
        x = _f64(x)
        f = _f64(prob.f if f is None else f)
        x0p = None if x0 is None else _dp(_f64(x0))
        rid = C.c_int64()
        rc = lib().tmg_server_submit(self.h, C.byref(pr), C.byref(mg), _PREC[precision], _dp(x), _dp(f), x0p,
                                     float(t_arrival), C.byref(rid))
        if rc != 0:
            raise TmgError(rc, self._err())
        self._meta[rid.value] = prob.nfree
        return rid.value

    def wait(self, rid, timeout=-1.0):
        st = C.c_int()
        rc = lib().tmg_server_wait(self.h, rid, float(timeout), C.byref(st))
        if rc < 0:
            raise TmgError(rc, self._err())
        return rc == 0, st.value

    def result(self, rid):
        n = self._meta[rid]
        u = np.zeros(n)
        it = C.c_int()
        rr = C.c_double()
        t = (C.c_double * 4)()
        rc = lib().tmg_server_result(self.h, rid, _dp(u), C.byref(it), C.byref(rr), t)
        if rc < 0:
            raise TmgError(rc, self._err())
        return {"u": u, "it": it.value, "relres": rr.value, "rc": rc, "t_arrival": t[0], "t_admit": t[1],
                "t_start": t[2], "t_done": t[3]}

    def release(self, rid):
        lib().tmg_server_release(self.h, rid)
        self._meta.pop(rid, None)

    def stats(self):
        s = (C.c_double * 8)()
        lib().tmg_server_stats(self.h, s)
        return {"bytes_reserved": s[0], "peak_bytes_reserved": s[1], "queued": int(s[2]), "running": int(s[3]),
                "done": int(s[4]), "clock": s[5], "peak_running": int(s[6]), "rebuilt": int(s[7])}

    def clock(self):
        return lib().tmg_server_clock(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().tmg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_workload(n=256, seed=4, sizes=(32, 48, 64, 80, 96, 112, 128), max_levels=5):
    """Requests of the serving benchmark: nelx ~ U{sizes}, nely = nelz = nelx/2,
    ngrids by the reference's level rules starting from max_levels, densities
    U[1e-3, 1] filtered (rmin 1.5) with seed 1000 + i, unit tip load Fy = -1 at
    (x = nelx, y ~ U{0..nely}, z ~ U{0..nelz})."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        nelx = int(rng.choice(sizes))
        nely = nelz = nelx // 2
        ty, tz = int(rng.integers(0, nely + 1)), int(rng.integers(0, nelz + 1))
        prob = P.cantilever(nelx, nely, nelz, load="point", tip=(ty, tz))
        L, opchol, ok = P.choose_ngrids(nelx, nely, nelz, max_levels)
        assert ok and opchol == 0
        x = P.random_density(nelx, nely, nelz, 1000 + i)
        reqs.append({"i": i, "prob": prob, "x": x, "ngrids": L})
    return reqs


def run_stream(server, reqs, rate, seed=4, precision="mixed"):
    """Submit reqs with exponential inter-arrival times (rate per second) in
    real time, wait for all, return per-request records."""
    rng = np.random.default_rng(seed + 1)
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    t_arr = np.cumsum(gaps) - gaps[0]
    base = server.clock() + 0.05
    ids = []

    def feeder():
        for r, ta in zip(reqs, t_arr):
            dt = base + ta - server.clock()
            if dt > 0:
                time.sleep(dt)
            ids.append(server.submit(r["prob"], r["x"], r["ngrids"], precision=precision, t_arrival=base + ta))

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    out = []
    for r, rid in zip(reqs, ids):
        ok, st = server.wait(rid)
        try:
            res = server.result(rid)
        except TmgError as e:
            # a request whose set-up or solve failed is reported, not fatal
            res = {"failed": True, "error": str(e)}
        res.update(i=r["i"], nfree=r["prob"].nfree, grid=(r["prob"].nelx, r["prob"].nely, r["prob"].nelz))
        server.release(rid)
        out.append(res)
    return out


def summarize(records):
    """Latency and throughput over the completed requests; failed ones are
    counted and listed (request index and error)."""
    failed = [r for r in records if r.get("failed")]
    done = [r for r in records if not r.get("failed")]
    lat = np.array([r["t_done"] - r["t_arrival"] for r in done])
    t_first = min(r["t_arrival"] for r in done)
    t_last = max(r["t_done"] for r in done)
    span = t_last - t_first
    dofit = sum(r["nfree"] * (r["it"] - 1) for r in done)
    out = {"requests": len(records), "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3), "max_ms": float(lat.max() * 1e3),
           "mean_ms": float(lat.mean() * 1e3), "makespan_s": span, "requests_per_s": len(done) / span,
           "dof_iters_per_s": dofit / span,
           "all_converged": not failed and all(r["rc"] == 0 and r["relres"] <= 1e-8 for r in done),
           "failed": len(failed)}
    if failed:
        out["failed_requests"] = [(r["i"], r["error"]) for r in failed[:8]]
    return out


def benchmark(reqs, budget_bytes=200e9, max_active=16, cap_sample=32, load=0.8, seed=4, precision="mixed",
              log=None):
    """Serving measurement: single-stream capacity from a serial run of the
    first ``cap_sample`` requests (one slot, all queued at once), then the
    whole stream replayed with Poisson arrivals at ``load`` x that capacity on
    ``max_active`` slots. Returns the stream summary with the capacity, rate
    and per-request latency percentiles."""
    srv = Server(budget_bytes, max_active=1)
    sample = reqs[:cap_sample]
    t0 = srv.clock()
    ids = [srv.submit(r["prob"], r["x"], r["ngrids"], precision=precision, t_arrival=t0) for r in sample]
    done = []
    for rid in ids:
        srv.wait(rid)
        done.append(srv.result(rid))
        srv.release(rid)
    span = max(d["t_done"] for d in done) - t0
    srv.close()
    cap = len(sample) / span
    serial = [d["t_done"] - d["t_admit"] for d in done]
    if log:
        log(f"serving: serial capacity {cap:.2f} req/s over {len(sample)} requests "
            f"(mean service {1e3 * float(np.mean(serial)):.1f} ms)")
    rate = load * cap
    srv = Server(budget_bytes, max_active=max_active)
    recs = run_stream(srv, reqs, rate, seed=seed, precision=precision)
    st = srv.stats()
    srv.close()
    out = summarize(recs)
    out.update(single_stream_capacity_rps=cap, arrival_rate_rps=rate, load=load, max_active=max_active,
               budget_bytes=budget_bytes, peak_bytes_reserved=st["peak_bytes_reserved"],
               rebuilt_after_failed_factorization=st["rebuilt"],
               mean_serial_service_ms=1e3 * float(np.mean(serial)))
    return out
