"""tmg_grad_noalloc (the compliance gradient on the library's static device
scratch, for callers that cannot allocate device memory) against tmg_grad on
the same handle and inputs, through the C-ABI: two calls in a row, the second
on a smaller problem, each bitwise equal to tmg_grad (the same kernels on the
same data); sizes beyond the scratch are rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

from tmg import problem as P

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _both(s, u_full, x, filt=None, freedens=None):
    from tmg._lib import lib
    u_full = np.ascontiguousarray(u_full, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    fd = None
    if freedens is not None:
        fdm = np.ascontiguousarray(freedens, dtype=np.uint8)
        fd = fdm.ctypes.data_as(C.POINTER(C.c_uint8))
    fh = filt.h if filt is not None else None
    g_ref = np.full(s.ndens, np.nan)
    g_na = np.full(s.ndens, np.nan)
    assert lib().tmg_grad(s.h, _dp(u_full), _dp(x), fd, fh, _dp(g_ref)) == 0
    assert lib().tmg_grad_noalloc(s.h, _dp(u_full), _dp(x), fd, fh, _dp(g_na)) == 0
    return g_ref, g_na


def test_grad_noalloc_matches_grad_twice():
    import tmg
    # first call: 64x32x32 with the density filter and a free-element mask
    _say("no-alloc gradient: 64x32x32 with filter and mask")
    prob = P.cantilever(64, 32, 32)
    x = P.random_density(64, 32, 32, 31)
    s = tmg.Solver(prob, 4, precision="mixed")
    s.update_density(x)
    u, it, rr, rc = s.solve(prob.f)
    assert rc == 0
    filt = tmg.Filter(64, 32, 32, 1.5, 2)
    mask = (np.random.default_rng(1).random(prob.nelem) > 0.1).astype(np.uint8)
    g_ref, g_na = _both(s, prob.full(u), x, filt, mask)
    assert np.all(np.isfinite(g_ref)) and np.array_equal(g_na, g_ref)
    # second call, smaller: a multiresolution 8x4x4 problem, no filter
    _say("no-alloc gradient: 8x4x4 multiresolution")
    prob2 = P.cantilever(8, 4, 4)
    xd = np.random.default_rng(2).uniform(0.1, 1.0, 8 * prob2.nelem)
    s2 = tmg.Solver(prob2, 2, precision="fp64", mr=2)
    s2.update_density(xd)
    u2, _, _, rc2 = s2.solve(prob2.f)
    assert rc2 == 0
    g_ref2, g_na2 = _both(s2, prob2.full(u2), xd)
    assert np.all(np.isfinite(g_ref2)) and np.array_equal(g_na2, g_ref2)
    # and the first problem again, after the smaller one used the scratch
    g_na3 = _both(s, prob.full(u), x, filt, mask)[1]
    assert np.array_equal(g_na3, g_ref)
    s.close()
    s2.close()


def test_grad_noalloc_rejects_what_does_not_fit():
    import tmg
    from tmg._lib import TMG_ERR_ARG, lib
    _say("no-alloc gradient: 128x64x64 multiresolution exceeds the scratch")
    prob = P.cantilever(128, 64, 64)
    s = tmg.Solver(prob, 5, precision="mixed", mr=2)
    xd = np.full(8 * prob.nelem, 0.5)
    u = np.zeros(prob.ndof)
    g = np.zeros(8 * prob.nelem)
    rc = lib().tmg_grad_noalloc(s.h, _dp(u), _dp(xd), None, None, _dp(g))
    assert rc == TMG_ERR_ARG
    assert "scratch" in lib().tmg_last_error(s.h).decode()
    s.close()
