"""C-ABI checks that need no GPU: the library loads, exports every function
include/tmg.h declares, and rejects invalid arguments before touching a device."""
import ctypes as C

import numpy as np
import pytest

import tmg
from tmg import _lib
from tmg import problem as P


def test_library_exports_every_declared_symbol():
    L = tmg.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 19
    for name in syms:
        assert hasattr(L, name), name
    assert set(syms) == set(_lib._SIGS), set(syms) ^ set(_lib._SIGS)


def _setup_rc(prob, ngrids=3, cycle=0, smoother=0, opchol=0, nu1=1, nu2=1, omega=0.6, prec=0, fd=None):
    fd = np.ascontiguousarray(prob.freedofs if fd is None else fd, dtype=np.int64)
    pr = _lib.Problem_t(prob.nelx, prob.nely, prob.nelz, prob.el, prob.eh, prob.ew, 1.0, prob.nu, 3.0, 1e-9,
                        fd.ctypes.data_as(C.POINTER(C.c_int64)), fd.size)
    mg = _lib.MgParams_t(ngrids, cycle, smoother, omega, nu1, nu2, 1e-8, 100, opchol, 60000)
    h = C.c_void_p()
    rc = tmg.lib().tmg_setup(C.byref(pr), C.byref(mg), prec, C.byref(h))
    return rc, _lib.last_error(None), h


@pytest.mark.parametrize("kw,code", [
    (dict(cycle=3), _lib.TMG_ERR_UNSUPPORTED),        # no fourth cycle type (0 V, 1 W, 2 FMV)
    (dict(smoother=2), _lib.TMG_ERR_UNSUPPORTED),     # SSOR
    (dict(opchol=1), _lib.TMG_ERR_UNSUPPORTED),       # PCG on the coarsest grid
    (dict(ngrids=1), _lib.TMG_ERR_ARG),
    (dict(ngrids=4), _lib.TMG_ERR_ARG),               # 4x... not divisible by 8
    (dict(omega=0.0), _lib.TMG_ERR_ARG),
    (dict(nu1=-1), _lib.TMG_ERR_ARG),
    (dict(prec=7), _lib.TMG_ERR_ARG),
])
def test_setup_rejects_bad_parameters(kw, code):
    prob = P.cantilever(8, 4, 4)
    rc, msg, h = _setup_rc(prob, **kw)
    assert rc == code and msg and not h.value


def test_setup_rejects_bad_freedofs():
    prob = P.cantilever(8, 4, 4)
    bad = prob.freedofs.copy()
    bad[3], bad[4] = bad[4], bad[3]        # not increasing
    rc, msg, _ = _setup_rc(prob, fd=bad)
    assert rc == _lib.TMG_ERR_ARG and "freedofs" in msg
    bad = prob.freedofs.copy()
    bad[-1] = prob.ndof + 1                # out of range
    rc, _, _ = _setup_rc(prob, fd=bad)
    assert rc == _lib.TMG_ERR_ARG


def test_filter_rejects_bad_arguments():
    h = C.c_void_p()
    assert tmg.lib().tmg_filter_create(4, 4, 4, 1.0, 1.0, 1.0, 1.5, 3, None, C.byref(h)) == _lib.TMG_ERR_ARG
    assert tmg.lib().tmg_filter_create(4, 4, 4, 1.0, 1.0, 1.0, -1.0, 2, None, C.byref(h)) == _lib.TMG_ERR_ARG


def test_null_handle_calls_fail_cleanly():
    L = tmg.lib()
    assert L.tmg_solve(None, None, None, None, None, None) == _lib.TMG_ERR_ARG
    assert L.tmg_update_density(None, None) == _lib.TMG_ERR_ARG
    assert L.tmg_ngrids(None) == 0
    L.tmg_destroy(None)


def test_server_rejects_bad_arguments_and_estimates():
    h = C.c_void_p()
    assert tmg.lib().tmg_server_create(0.0, 4, C.byref(h)) == _lib.TMG_ERR_ARG
    assert tmg.lib().tmg_server_create(1e9, 0, C.byref(h)) == _lib.TMG_ERR_ARG
    from tmg import serve
    small = serve.estimate_bytes(P.cantilever(32, 16, 16), 5)
    large = serve.estimate_bytes(P.cantilever(128, 64, 64), 5)
    assert 0 < small < large < 2e9
    assert 0 < serve.estimate_bytes(P.cantilever(128, 64, 64), 5, "fp64") < 2e9


def test_serving_summary_counts_failed_requests():
    """A request that fails (set-up or solve error) is counted and listed by
    the serving summary; latency and throughput cover the completed ones."""
    from tmg import serve
    ok = [{"t_arrival": 0.1 * k, "t_done": 0.1 * k + 0.02, "nfree": 1000, "it": 11, "rc": 0, "relres": 5e-9, "i": k}
          for k in range(9)]
    bad = {"failed": True, "error": "tmg error -4: not positive definite", "i": 9, "nfree": 1000}
    out = serve.summarize(ok + [bad])
    assert out["requests"] == 10 and out["failed"] == 1
    assert out["failed_requests"] == [(9, "tmg error -4: not positive definite")]
    assert not out["all_converged"]
    assert abs(out["p50_ms"] - 20.0) < 1e-9
    out2 = serve.summarize(ok)
    assert out2["failed"] == 0 and out2["all_converged"] and "failed_requests" not in out2
