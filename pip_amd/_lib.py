"""ctypes binding of libpipgpu.so (the C-ABI in include/pipgpu.h).

There is no fallback: if the library is missing or cannot be loaded, every
entry point raises.  Build it with ``python -c "import __graft_entry__ as g;
g.build()"`` (or ``python -m pip_amd.build``).
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PIPGPU_LIB") or os.path.join(HERE, "lib", "libpipgpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pipgpu.h")

PIP_OK = 0


class PipCfg(ctypes.Structure):
    """pip_cfg of include/pipgpu.h."""
    _fields_ = [
        ("ndim", ctypes.c_int), ("ix", ctypes.c_int), ("jx", ctypes.c_int),
        ("nvar_m", ctypes.c_int), ("flag_divb", ctypes.c_int), ("t_order", ctypes.c_int),
        ("flag_ir", ctypes.c_int), ("flag_ir_type", ctypes.c_int), ("flag_col", ctypes.c_int),
        ("ac_field", ctypes.c_int), ("flag_bnd", ctypes.c_int * 4),
        ("gm", ctypes.c_double), ("safety", ctypes.c_double), ("ro_lim", ctypes.c_double),
        ("pr_lim", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("col", ctypes.c_double), ("db_clean", ctypes.c_double), ("t_ir", ctypes.c_double),
        ("T0", ctypes.c_double), ("n0", ctypes.c_double), ("beta", ctypes.c_double),
        ("flag_pip_imp", ctypes.c_int), ("pip_imp_factor", ctypes.c_double),
    ]


def cfg_to_c(cfg) -> PipCfg:
    c = PipCfg()
    c.ndim, c.ix, c.jx, c.nvar_m = cfg.ndim, cfg.ix, cfg.jx, cfg.nvar_m
    c.flag_divb, c.t_order, c.flag_ir = cfg.flag_divb, cfg.t_order, cfg.flag_ir
    c.flag_ir_type, c.flag_col, c.ac_field = cfg.flag_ir_type, cfg.flag_col, int(cfg.ac_field)
    for i in range(4):
        c.flag_bnd[i] = cfg.flag_bnd[i]
    c.gm, c.safety, c.ro_lim, c.pr_lim = cfg.gm, cfg.safety, cfg.ro_lim, cfg.pr_lim
    c.dx, c.dy, c.col, c.db_clean = cfg.dx, cfg.dy, cfg.col, cfg.db_clean
    c.t_ir, c.T0, c.n0, c.beta = cfg.t_ir, cfg.T0, cfg.n0, cfg.beta
    c.flag_pip_imp, c.pip_imp_factor = cfg.flag_pip_imp, cfg.pip_imp_factor
    return c


class PipEnsMember(ctypes.Structure):
    """pip_ens_member of include/pipgpu.h."""
    _fields_ = [("nx", ctypes.c_int), ("dx", ctypes.c_double), ("col", ctypes.c_double),
                ("tend", ctypes.c_double), ("max_steps", ctypes.c_int), ("cost", ctypes.c_double),
                ("ny", ctypes.c_int), ("dy", ctypes.c_double)]


class PipError(RuntimeError):
    pass


_lib = None
_P = ctypes.POINTER(ctypes.c_double)
_V = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PI = ctypes.POINTER(ctypes.c_int)
_PD = ctypes.POINTER(ctypes.c_double)
_PL = ctypes.POINTER(ctypes.c_long)
_PP = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "pip_create": [ctypes.POINTER(PipCfg), _PP],
    "pip_destroy": [_V],
    "pip_upload": [_V, _P, _P, _P],
    "pip_download": [_V, _P, _P],
    "pip_prologue": [_V, _PD],
    "pip_set_ir": [_V],
    "pip_cfl": [_V, _PD, _I],
    "pip_run_solver": [_V],
    "pip_bnd": [_V],
    "pip_advance": [_V, _I, _D, _PI, _PD],
    "pip_enqueue_steps": [_V, _I, _D],
    "pip_sync": [_V],
    "pip_get_scalars": [_V, _PD, _PD, _PD, _PL],
    "pip_set_time": [_V, _D, _D],
    "pip_status": [_V, _PI],
    "pip_get_rates": [_V, _P, _P, _P],
    "pip_get_ac": [_V, _P],
    "pip_set_arb_heat": [_V, _P],
    "pip_device_state": [_V, _PP, _PP],
    "pip_get_stream": [_V, _PP],
    "pip_set_tile": [_V, _I, _I],
    "pip_set_xcd_remap": [_V, _I],
    "pip_set_two_pass": [_V, _I],
    "pip_set_arith": [_V, _I],
    "pip_set_step_graph": [_V, _I],
    "pip_time_steps": [_V, _I, _PD, _PD, _PD],
    "pip_set_state_device": [_V, _V],
    "pip_get_state_device": [_V, _V],
    "pip_set_ac_device": [_V, _V],
    "pip_record_steps": [_V, _I, _V],
    "pip_replay_steps": [_V, _I, _V, _V, _V, _V, _V],
    "pip_record_steps_saved": [_V, _I, _V, _V, _V, _V, _V],
    "pip_adj_seed": [_V, _V],
    "pip_adj_step": [_V, _V, _V, _V, _V, _V, _V, _V],
    "pip_adj_step_noalloc": [_V, _V, _V, _V, _V, _V, _V, _V],
    "pip_ens_create": [ctypes.POINTER(PipCfg), ctypes.POINTER(PipEnsMember), _I, _PP],
    "pip_ens_upload": [_V, _I, _P, _P],
    "pip_ens_download": [_V, _I, _P, _P],
    "pip_ens_set_tile": [_V, _I],
    "pip_ens_set_mode": [_V, _I],
    "pip_ens_reset": [_V],
    "pip_ens_run": [_V, _I, _PD],
    "pip_ens_result": [_V, _I, _PD, _PL, _PI],
    "pip_ens_destroy": [_V],
    "pip_check_arith": [_P, _P, _P, _P, ctypes.c_longlong],
}


def header_functions():
    """Function names declared in include/pipgpu.h."""
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(pip_\w+)\s*\(", src, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PipError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.pip_last_error.restype = ctypes.c_char_p
        L.pip_last_error.argtypes = []
        L.pip_version.restype = ctypes.c_char_p
        L.pip_version.argtypes = []
        _lib = L
    return _lib


def check(rc: int):
    if rc != PIP_OK:
        raise PipError(f"pipgpu error {rc}: {lib().pip_last_error().decode()}")
    return rc
