"""ctypes binding of libtmg.so (the C-ABI in include/tmg.h).

The library is built in-tree by ``csrc/Makefile``. There is no fallback: if
the shared library is missing or fails to load, importing the solver raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtmg.so")
HEADER_PATH = os.path.abspath(os.path.join(_HERE, "..", "..", "include", "tmg.h"))

TMG_FP64 = 0
TMG_MIXED = 1

TMG_OK = 0
TMG_NOT_CONVERGED = 1
TMG_ERR_ARG = -1
TMG_ERR_UNSUPPORTED = -2
TMG_ERR_OOM = -3
TMG_ERR_NOT_SPD = -4
TMG_ERR_HIP = -5
TMG_ERR_STATE = -6
TMG_ERR_BUDGET = -7


class TmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tmg error {code}: {msg}")
        self.code = code


class Problem_t(C.Structure):
    _fields_ = [
        ("nelx", C.c_int), ("nely", C.c_int), ("nelz", C.c_int),
        ("el", C.c_double), ("eh", C.c_double), ("ew", C.c_double),
        ("E0", C.c_double), ("nu", C.c_double), ("penal", C.c_double), ("emin", C.c_double),
        ("freedofs", C.POINTER(C.c_int64)), ("nfree", C.c_int64),
    ]


class MgParams_t(C.Structure):
    _fields_ = [
        ("ngrids", C.c_int), ("cycle", C.c_int), ("smoother", C.c_int), ("omega", C.c_double),
        ("smoothiter1", C.c_int), ("smoothiter2", C.c_int), ("tolMG", C.c_double),
        ("maxiterMG", C.c_int), ("opchol", C.c_int), ("nmaxchol", C.c_int),
    ]


_lib = None

_p = C.c_void_p
_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pi64 = C.POINTER(C.c_int64)
_pu8 = C.POINTER(C.c_uint8)

_SIGS = {
    "tmg_setup": (_i, [C.POINTER(Problem_t), C.POINTER(MgParams_t), _i, C.POINTER(_p)]),
    "tmg_setup_mr": (_i, [C.POINTER(Problem_t), C.POINTER(MgParams_t), _i, _i, C.POINTER(_p)]),
    "tmg_update_density": (_i, [_p, _pd]),
    "tmg_update_density_device": (_i, [_p, _p]),
    "tmg_solve": (_i, [_p, _pd, _pd, _pd, _pi, _pd]),
    "tmg_solve_device": (_i, [_p, _p, _p, _pi, _pd]),
    "tmg_vcycle": (_i, [_p, _pd, _pd]),
    "tmg_smooth": (_i, [_p, _i, _pd, _pd, _i, _pd]),
    "tmg_apply": (_i, [_p, _i, _pd, _pd]),
    "tmg_prolong": (_i, [_p, _i, _pd, _pd]),
    "tmg_restrict": (_i, [_p, _i, _pd, _pd]),
    "tmg_level_info": (_i, [_p, _i, _pi, _pi64, _pi64]),
    "tmg_ngrids": (_i, [_p]),
    "tmg_solve_stats": (_i, [_p, _pi, _pi]),
    "tmg_grad": (_i, [_p, _pd, _pd, _pu8, _p, _pd]),
    "tmg_grad_noalloc": (_i, [_p, _pd, _pd, _pu8, _p, _pd]),
    "tmg_filter_create": (_i, [_i, _i, _i, _d, _d, _d, _d, _i, _pu8, C.POINTER(_p)]),
    "tmg_filter_apply": (_i, [_p, _pd, _i, _pd]),
    "tmg_filter_rowsum": (_i, [_p, _pd]),
    "tmg_filter_destroy": (None, [_p]),
    "tmg_time_kernel": (_i, [_p, _i, _i, _pd]),
    "tmg_time_in_solve": (_i, [_p, _i, _pd, _pi]),
    "tmg_device_bytes": (C.c_int64, [_p]),
    "tmg_lp_volume": (_i, [C.c_int64, _pd, _pd, _pd, _pd, _d, _i, _pd, _pd]),
    "tmg_server_create": (_i, [_d, _i, C.POINTER(_p)]),
    "tmg_server_submit": (_i, [_p, C.POINTER(Problem_t), C.POINTER(MgParams_t), _i, _pd, _pd, _pd, _d, _pi64]),
    "tmg_server_wait": (_i, [_p, C.c_int64, _d, _pi]),
    "tmg_server_result": (_i, [_p, C.c_int64, _pd, _pi, _pd, _pd]),
    "tmg_server_release": (_i, [_p, C.c_int64]),
    "tmg_server_stats": (_i, [_p, _pd]),
    "tmg_server_clock": (_d, [_p]),
    "tmg_server_estimate": (C.c_int64, [C.POINTER(Problem_t), C.POINTER(MgParams_t), _i]),
    "tmg_server_last_error": (C.c_char_p, [_p]),
    "tmg_server_destroy": (None, [_p]),
    "tmg_last_error": (C.c_char_p, [_p]),
    "tmg_destroy": (None, [_p]),
}


def header_symbols():
    """Names of the functions include/tmg.h declares."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tmg_[a-z_0-9]+)\s*\(", txt)))


def lib():
    """Load libtmg.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; build it with __graft_entry__.build() "
                              "or `make -C topopt-slp-3d_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error(h=None) -> str:
    m = lib().tmg_last_error(h)
    return m.decode() if m else ""


def check(rc, h=None, allow_not_converged=False):
    if rc < 0 or (rc > 0 and not allow_not_converged):
        raise TmgError(rc, last_error(h))
    return rc
