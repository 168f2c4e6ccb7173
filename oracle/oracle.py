"""ctypes wrapper of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check / time the HIP path; the product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liborc.so")


class OrcCfg(ctypes.Structure):
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


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcCfg)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_set_state", "orc_get_state"):
            getattr(L, name).argtypes = [ctypes.c_void_p, P, P]
        for name in ("orc_set_ac", "orc_get_ac"):
            getattr(L, name).argtypes = [ctypes.c_void_p, P]
        L.orc_get_rates.argtypes = [ctypes.c_void_p, P, P, P, P]
        L.orc_set_rates.argtypes = [ctypes.c_void_p, P, P]
        L.orc_set_arb_heat.argtypes = [ctypes.c_void_p, P]
        L.orc_get_scalars.argtypes = [ctypes.c_void_p, P]
        L.orc_set_scalars.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_set_stage_save.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        for name in ("orc_pipbnd", "orc_set_ir", "orc_run_solver", "orc_prologue"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_set_coefficients.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_fixed_rates.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_cfl.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_cfl.restype = ctypes.c_double
        L.orc_step.argtypes = [ctypes.c_void_p]
        L.orc_step.restype = ctypes.c_int
        L.orc_stop_sim.argtypes = [ctypes.c_void_p]
        L.orc_stop_sim.restype = ctypes.c_int
        for name in ("orc_exp", "orc_log"):
            getattr(L, name).argtypes = [ctypes.c_double]
            getattr(L, name).restype = ctypes.c_double
        L.orc_pow.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_pow.restype = ctypes.c_double
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_minmod.argtypes = [ctypes.c_double] * 3
        L.orc_minmod.restype = ctypes.c_double
        L.orc_hllc_flux.argtypes = [P, P, ctypes.c_double, P]
        L.orc_hlld_flux.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, P]
        L.orc_set_flux.argtypes = [ctypes.c_void_p, P, P]
        L.orc_source_pip.argtypes = [ctypes.c_void_p, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_cfg(cfg) -> OrcCfg:
    c = OrcCfg()
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


class Oracle:
    """One reference-path run on the CPU (single rank)."""

    def __init__(self, cfg, U_h, U_m, ac=None, prologue=True):
        self.cfg = cfg
        self._c = make_cfg(cfg)
        L = lib()
        self.h = L.orc_create(ctypes.byref(self._c))
        uh = np.ascontiguousarray(U_h, dtype=np.float64)
        um = np.ascontiguousarray(U_m, dtype=np.float64)
        L.orc_set_state(self.h, _p(uh), _p(um))
        if ac is not None:
            L.orc_set_ac(self.h, _p(np.ascontiguousarray(ac, dtype=np.float64)))
        if prologue:
            L.orc_prologue(self.h)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # state ---------------------------------------------------------------
    def state(self):
        c = self.cfg
        uh = np.empty((5, c.jx, c.ix))
        um = np.empty((c.nvar_m, c.jx, c.ix))
        lib().orc_get_state(self.h, _p(uh), _p(um))
        return uh, um

    def set_state(self, U_h, U_m):
        lib().orc_set_state(self.h, _p(np.ascontiguousarray(U_h, dtype=np.float64)),
                            _p(np.ascontiguousarray(U_m, dtype=np.float64)))

    def rates(self):
        c = self.cfg
        out = [np.empty((c.jx, c.ix)) for _ in range(4)]
        lib().orc_get_rates(self.h, *[_p(o) for o in out])
        return out

    def set_rates(self, gm_rec, gm_ion):
        lib().orc_set_rates(self.h, _p(np.ascontiguousarray(gm_rec, dtype=np.float64)),
                            _p(np.ascontiguousarray(gm_ion, dtype=np.float64)))

    def set_arb_heat(self, arb_heat):
        lib().orc_set_arb_heat(self.h, _p(np.ascontiguousarray(arb_heat, dtype=np.float64)))

    def fixed_rates(self, on=True):
        lib().orc_set_fixed_rates(self.h, int(on))

    def scalars(self):
        o = np.empty(4)
        lib().orc_get_scalars(self.h, _p(o))
        return dict(dt=o[0], time=o[1], cmax=o[2], nt=int(o[3]))

    def set_scalars(self, dt: float, cmax: float):
        """Test hook: dt and GLM speed cmax for the next run_solver."""
        lib().orc_set_scalars(self.h, float(dt), float(cmax))

    def run_solver_stage1(self):
        """run_solver, also returning the stage-1 result (nvar_m+5, jx, ix)
        as (U_m planes, U_h planes)."""
        c = self.cfg
        buf = np.empty((c.nvar_m + 5, c.jx, c.ix))
        lib().orc_set_stage_save(self.h, buf.ctypes.data)
        try:
            lib().orc_run_solver(self.h)
        finally:
            lib().orc_set_stage_save(self.h, None)
        return buf

    # reference entry points ---------------------------------------------
    def step(self) -> int:
        return lib().orc_step(self.h)

    def run(self, nsteps=None, tend=None):
        """main.f90:31-68: loop while time < tend (or for nsteps)."""
        n = 0
        while True:
            if nsteps is not None and n >= nsteps:
                break
            if tend is not None and self.scalars()["time"] >= tend:
                break
            st = self.step()
            n += 1
            if st:
                return st
        return 0

    def cfl(self, into_global=True) -> float:
        return lib().orc_cfl(self.h, int(into_global))

    def run_solver(self):
        lib().orc_run_solver(self.h)

    def set_ir(self):
        lib().orc_set_ir(self.h)

    def pipbnd(self):
        lib().orc_pipbnd(self.h)

    def flux(self):
        c = self.cfg
        N = c.ix * c.jx
        fh = np.zeros((2, 5, c.jx, c.ix))
        fm = np.zeros((2, c.nvar_m, c.jx, c.ix))
        lib().orc_set_flux(self.h, _p(fh), _p(fm))
        return fh, fm

    def source(self):
        c = self.cfg
        sh = np.zeros((5, c.jx, c.ix))
        sm = np.zeros((c.nvar_m, c.jx, c.ix))
        lib().orc_source_pip(self.h, _p(sh), _p(sm))
        return sh, sm


def set_threads(n: int) -> int:
    """Threads of the oracle's elementwise loops (1 = serial); returns the
    previous setting.  Results do not depend on it."""
    return lib().orc_set_threads(int(n))


def minmod(a, b, c) -> float:
    return lib().orc_minmod(a, b, c)


def hllc_flux(vl, vr, gm=5.0 / 3.0):
    vl = np.ascontiguousarray(vl, dtype=np.float64)
    vr = np.ascontiguousarray(vr, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_hllc_flux(_p(vl), _p(vr), gm, _p(out))
    return out


def hlld_flux(vl, vr, gm=5.0 / 3.0, divb=0, cmax=0.0):
    vl = np.ascontiguousarray(vl, dtype=np.float64)
    vr = np.ascontiguousarray(vr, dtype=np.float64)
    nv = len(vl)
    out = np.zeros(9)
    lib().orc_hlld_flux(_p(vl), _p(vr), nv, divb, cmax, gm, _p(out))
    return out[:nv]
