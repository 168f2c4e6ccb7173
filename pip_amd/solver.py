"""Host-side mirror of the reference's time-loop interface over libpipgpu.

``PIPSolver`` owns one device context.  Its methods carry the reference's
procedure names and meaning:

  prologue()    Initial_rot.f90:140-160 (ghost fill, set_coefficients, cfl)
  set_ir()      PIP_rot.f90:127-311  (case 3 rates, rewrites E)
  cfl()         Scheme_rot.f90:183-213 (rewrites U, returns dt)
  run_solver()  Solver_rot.f90:42-81  (SSP-RK stages, sources, boundaries)
  step()        one pass of main.f90:31-62 through the calls above
  advance()     the same loop fused on the device (no host round trips)

The state stays on the GPU; ``state()`` downloads it in the reference
layout (nvar, jx, ix).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


ARITH = {"exact": 0, "relaxed": 1}  # PIP_ARITH_EXACT, PIP_ARITH_RELAXED


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PIPSolver:
    def __init__(self, cfg, U_h, U_m, ac=None, prologue=True, tile=None, arith="exact"):
        """arith: "exact" (the reference's arithmetic, bitwise the CPU path)
        or "relaxed" (FMA contraction, reciprocal-multiply quotients,
        hardware min/max; see pip_set_arith)."""
        self.cfg = cfg
        self._ccfg = _lib.cfg_to_c(cfg)
        h = ctypes.c_void_p()
        check(lib().pip_create(ctypes.byref(self._ccfg), ctypes.byref(h)))
        self.h = h
        self.time_host = 0.0
        if tile is not None:
            self.set_tile(*tile)  # (tx, ty[, xcd_remap[, two_pass stage mask]])
        self.set_arith(arith)
        self.upload(U_h, U_m, ac)
        if prologue:
            self.prologue()

    def close(self):
        if getattr(self, "h", None):
            lib().pip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- data
    def upload(self, U_h, U_m, ac=None):
        c = self.cfg
        uh = np.ascontiguousarray(U_h, dtype=np.float64).reshape(5, c.jx, c.ix)
        um = np.ascontiguousarray(U_m, dtype=np.float64).reshape(c.nvar_m, c.jx, c.ix)
        if c.ac_field:
            if ac is None:
                raise ValueError("cfg.ac_field needs an ac array")
            a = np.ascontiguousarray(ac, dtype=np.float64).reshape(c.jx, c.ix)
            check(lib().pip_upload(self.h, _p(uh), _p(um), _p(a)))
        else:
            check(lib().pip_upload(self.h, _p(uh), _p(um), None))

    def state(self):
        c = self.cfg
        uh = np.empty((5, c.jx, c.ix))
        um = np.empty((c.nvar_m, c.jx, c.ix))
        check(lib().pip_download(self.h, _p(uh), _p(um)))
        return uh, um

    def scalars(self):
        dt, t, cm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nt = ctypes.c_long()
        check(lib().pip_get_scalars(self.h, ctypes.byref(dt), ctypes.byref(t), ctypes.byref(cm),
                                    ctypes.byref(nt)))
        return dict(dt=dt.value, time=t.value, cmax=cm.value, nt=nt.value)

    def status(self) -> int:
        f = ctypes.c_int()
        check(lib().pip_status(self.h, ctypes.byref(f)))
        return f.value

    def ac_field(self):
        """The collision coefficient field as the output routines see it."""
        c = self.cfg
        out = np.empty((c.jx, c.ix))
        check(lib().pip_get_ac(self.h, _p(out)))
        return out

    def set_arb_heat(self, arb_heat):
        """Upload arb_heat for a restart (kept by the next prologue)."""
        c = self.cfg
        a = np.ascontiguousarray(arb_heat, dtype=np.float64).reshape(c.jx, c.ix)
        check(lib().pip_set_arb_heat(self.h, _p(a)))

    def set_time(self, time: float, dt: float | None = None):
        """Set the simulation time (restart) and, optionally, dt."""
        if dt is None:
            dt = self.scalars()["dt"]
        check(lib().pip_set_time(self.h, float(time), float(dt)))
        self.time_host = float(time)

    def rates(self):
        c = self.cfg
        out = [np.empty((c.jx, c.ix)) for _ in range(3)]
        check(lib().pip_get_rates(self.h, *[_p(o) for o in out]))
        return out

    # ------------------------------------------------ reference procedures
    def prologue(self) -> float:
        dt = ctypes.c_double()
        check(lib().pip_prologue(self.h, ctypes.byref(dt)))
        self.time_host = 0.0
        return dt.value

    def set_ir(self):
        check(lib().pip_set_ir(self.h))

    def cfl(self, into_global=True) -> float:
        dt = ctypes.c_double()
        check(lib().pip_cfl(self.h, ctypes.byref(dt), int(into_global)))
        return dt.value

    def run_solver(self):
        check(lib().pip_run_solver(self.h))

    def pipbnd(self):
        check(lib().pip_bnd(self.h))

    def step(self) -> float:
        """main.f90:31-58 through the individual procedures (host loop)."""
        c = self.cfg
        if c.flag_ir >= 1:
            self.set_ir()
        dt = self.cfl(True)
        self.run_solver()
        if c.flag_ir >= 1:
            self.set_ir()
        self.cfl(False)
        self.time_host = self.time_host + dt
        return dt

    # ------------------------------------------------------- fused loop
    def advance(self, nsteps: int | None = None, tend: float | None = None):
        n = ctypes.c_int()
        t = ctypes.c_double()
        check(lib().pip_advance(self.h, int(nsteps if nsteps is not None else 2 ** 30),
                                float(tend if tend is not None else 1.0e300),
                                ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def enqueue(self, nsteps: int, tend: float = 1.0e300):
        check(lib().pip_enqueue_steps(self.h, int(nsteps), float(tend)))

    def sync(self):
        check(lib().pip_sync(self.h))

    def set_tile(self, tx: int, ty: int, xcd_remap: int | None = None, two_pass: int | None = None):
        """Tile shape, XCD tile order and the mask of stages that take the
        fluids one at a time (bit s-1: stage s)."""
        check(lib().pip_set_tile(self.h, int(tx), int(ty)))
        if xcd_remap is not None:
            check(lib().pip_set_xcd_remap(self.h, int(xcd_remap)))
        if two_pass is not None:
            check(lib().pip_set_two_pass(self.h, int(two_pass)))

    def set_arith(self, arith: str):
        check(lib().pip_set_arith(self.h, ARITH[arith]))
        self.arith = arith

    def time_steps(self, nsteps: int):
        tt, t1, t2 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().pip_time_steps(self.h, int(nsteps), ctypes.byref(tt), ctypes.byref(t1),
                                   ctypes.byref(t2)))
        return dict(total_ms=tt.value, stage1_ms=t1.value, stage2_ms=t2.value)

    def device_state(self):
        U = ctypes.c_void_p()
        ac = ctypes.c_void_p()
        check(lib().pip_device_state(self.h, ctypes.byref(U), ctypes.byref(ac)))
        return U.value, ac.value

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().pip_get_stream(self.h, ctypes.byref(s)))
        return s.value
