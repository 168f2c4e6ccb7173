"""Batched ensembles of independent 1-D two-fluid runs (the pip_ens_* entry
points of include/pipgpu.h): one persistent launch, host-driven lockstep
launches, or lockstep chunks captured as a hipGraph.

Each member is a full run of the reference time loop (main.f90:31-68) with
its own size, coefficients, dt sequence and end time.  Members are scheduled
longest-first from a device queue; the launch can be captured once in a
hipGraph and replayed.  Results equal the plain per-member runs bit for bit.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, lib


MODES = {"persistent": 0, "lockstep": 1, "graph": 2}  # pip_ens_set_mode


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cost_estimate(cfg, U_h, U_m, tend: float) -> float:
    """Expected work nx * steps, with steps from the CFL and collision limits
    of the initial state (Scheme_rot.f90:245-313)."""
    gm = cfg.gm
    um, uh = U_m[:, 0, 2:-2], U_h[:, 0, 2:-2]
    rho = um[0]
    b2 = um[5] ** 2 + um[6] ** 2 + um[7] ** 2
    v = np.abs(um[1] / rho)
    p = (gm - 1) * (um[4] - 0.5 * (um[1] ** 2) / rho - 0.5 * b2)
    cf = np.sqrt(np.maximum(gm * p / rho + b2 / rho, 0.0))
    dt = cfg.safety * cfg.dx / np.max(v + cf)
    dt = min(dt, min(cfg.safety, 0.3) / max(np.max(rho * cfg.col), np.max(uh[0] * cfg.col)))
    if not (np.isfinite(dt) and dt > 0.0):
        return float(cfg.nx)  # a state the first step's stop_sim ends: one step's work
    return cfg.nx * min(math.ceil(tend / dt), 1 << 40)


class Ensemble:
    def __init__(self, members, tend: float, max_steps: int = 0, tile: int = 256):
        """members: list of (cfg, U_h, U_m) with configurations of one
        dimension that share gm, safety, limits and boundary flags: 1-D, or
        2-D with psi (flag_divb = 1; lockstep and graph modes)."""
        self.cfgs = [m[0] for m in members]
        base = self.cfgs[0]
        if any(c.ndim != base.ndim or c.nvar_m != base.nvar_m for c in self.cfgs):
            raise ValueError("ensemble members must share ndim and nvar_m")
        self._base = _lib.cfg_to_c(base)
        arr = (_lib.PipEnsMember * len(members))()
        for k, (cfg, uh, um) in enumerate(members):
            arr[k].nx = cfg.nx
            arr[k].dx = cfg.dx
            arr[k].col = cfg.col
            arr[k].tend = tend
            arr[k].max_steps = max_steps
            arr[k].cost = float(cost_estimate(cfg, uh, um, tend)) if cfg.ndim == 1 else float(cfg.nx * cfg.ny)
            arr[k].ny = cfg.ny if cfg.ndim == 2 else 0
            arr[k].dy = cfg.dy if cfg.ndim == 2 else 0.0
        self._arr = arr
        h = ctypes.c_void_p()
        check(lib().pip_ens_create(ctypes.byref(self._base), arr, len(members), ctypes.byref(h)))
        self.h = h
        for k, (cfg, uh, um) in enumerate(members):
            check(lib().pip_ens_upload(self.h, k, _p(np.ascontiguousarray(uh)), _p(np.ascontiguousarray(um))))
        self.set_tile(tile)
        self.mode = "persistent"
        if base.ndim == 2:
            self.set_mode("lockstep")

    def close(self):
        if getattr(self, "h", None):
            lib().pip_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, tx: int):
        check(lib().pip_ens_set_tile(self.h, int(tx)))
        self.tile = tx

    def set_mode(self, mode: str):
        """'persistent': one launch, a workgroup per member (hipGraph replay
        with graph=True); 'lockstep': every member's tiles spread over
        workgroups, two stage launches and a step end per step, host-driven
        64-step chunks; 'graph': the lockstep chunk captured once as a
        hipGraph over every member's tiles and replayed until all members
        stop."""
        check(lib().pip_ens_set_mode(self.h, MODES[mode]))
        self.mode = mode

    def reset(self):
        check(lib().pip_ens_reset(self.h))

    def run(self, graph: bool = True) -> float:
        ms = ctypes.c_double()
        check(lib().pip_ens_run(self.h, int(graph), ctypes.byref(ms)))
        return ms.value

    def result(self, m: int):
        t, nt, st = ctypes.c_double(), ctypes.c_long(), ctypes.c_int()
        check(lib().pip_ens_result(self.h, m, ctypes.byref(t), ctypes.byref(nt), ctypes.byref(st)))
        return dict(time=t.value, nt=nt.value, status=st.value)

    def state(self, m: int):
        cfg = self.cfgs[m]
        uh = np.empty((5, cfg.jx, cfg.ix))
        um = np.empty((cfg.nvar_m, cfg.jx, cfg.ix))
        check(lib().pip_ens_download(self.h, m, _p(uh), _p(um)))
        return uh, um
