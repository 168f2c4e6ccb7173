"""Snapshots in the reference's HDF5 output layout (t####.h5) through
libpipio.so (include/pipio.h).

The reference's output routine (IO_rot.f90:39-110) writes, per output, the
coordinate arrays of set_coordinate, the time and every state plane under
the names of set_initial_out (ro_p ... bz, ps, ro_n ... en_n), plus the
coefficient fields def_varfiles selects; its restart reads the same names.
``write`` / ``read`` do this for host arrays, ``write_solver`` /
``restart_solver`` for a PIPSolver (state downloaded from / uploaded to the
device).  There is no fallback: a missing libpipio.so raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build
from .config import MARGIN, PIPConfig

_P = ctypes.POINTER(ctypes.c_double)
_io = None


class PipioLayout(ctypes.Structure):
    """pipio_layout of include/pipio.h."""
    _fields_ = [(n, ctypes.c_int) for n in (
        "ix", "jx", "kx", "nvar_m", "flag_divb", "flag_col", "flag_ir", "flag_ir_type",
        "ac_sav", "xi_sav", "col_sav", "ion_sav", "rec_sav", "heat_sav", "ps_sav")]


class PipioFields(ctypes.Structure):
    """pipio_fields of include/pipio.h."""
    _fields_ = [("time", ctypes.c_double)] + [(n, _P) for n in (
        "x", "y", "z", "dx", "dy", "dz", "U_h", "U_m", "ac", "xi", "gm_ion", "gm_rec", "arb_heat")]


class PipioError(RuntimeError):
    pass


def lib():
    global _io
    if _io is None:
        if not os.path.exists(build.IOLIB):
            raise PipioError(f"{build.IOLIB} is not built (needs the HDF5 C library); run "
                             "__graft_entry__.build()")
        L = ctypes.CDLL(build.IOLIB)
        L.pipio_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(PipioLayout),
                                  ctypes.POINTER(PipioFields)]
        L.pipio_write.restype = ctypes.c_int
        L.pipio_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(PipioLayout)] + [_P] * 13
        L.pipio_read.restype = ctypes.c_int
        L.pipio_last_error.restype = ctypes.c_char_p
        L.pipio_last_error.argtypes = []
        _io = L
    return _io


def _check(rc):
    if rc != 0:
        raise PipioError(f"pipio error {rc}: {lib().pipio_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def grid_arrays(cfg: PIPConfig):
    """x, y, z, dx, dy, dz of set_coordinate (Model_rot.f90:394-458) for a
    uniform single-rank grid: the first interior centre at start + dr/2,
    then x(i) = x(i-1) + dx(i-1) upwards and x(i) = x(i+1) - dx(i)
    downwards; unused directions have dr = 0 and d = 100 in their first
    entry."""
    def axis(n, margin, start, end):
        dr = (end - start) / (n - 2 * margin)
        d = np.full(n, dr)
        c = np.empty(n)
        z0 = margin  # zeros(:) = margin + 1 (1-based)
        c[z0] = start + 0.5 * dr
        for i in range(z0 + 1, n):
            c[i] = c[i - 1] + d[i - 1]
        for i in range(z0 - 1, -1, -1):
            c[i] = c[i + 1] - d[i]
        return c, d
    x, dx = axis(cfg.ix, MARGIN, cfg.xmin, cfg.xmax)
    if cfg.ndim >= 2:
        y, dy = axis(cfg.jx, MARGIN, cfg.ymin, cfg.ymax)
    else:
        y, dy = axis(1, 0, 0.0, 0.0)
        dy[0] = 100.0
    z, dz = axis(1, 0, 0.0, 0.0)
    dz[0] = 100.0
    return x, y, z, dx, dy, dz


def layout(cfg: PIPConfig, **sav) -> PipioLayout:
    """The datasets of one output; sav: the setting.txt *_sav keys (0 = write)."""
    L = PipioLayout()
    L.ix, L.jx, L.kx, L.nvar_m = cfg.ix, cfg.jx, 1, cfg.nvar_m
    L.flag_divb, L.flag_col, L.flag_ir, L.flag_ir_type = cfg.flag_divb, cfg.flag_col, cfg.flag_ir, cfg.flag_ir_type
    for k in ("ac_sav", "xi_sav", "col_sav", "ion_sav", "rec_sav", "heat_sav", "ps_sav"):
        setattr(L, k, int(sav.pop(k, 0)))
    if sav:
        raise ValueError(f"unknown save keys {sorted(sav)}")
    return L


def _plane(cfg, a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(cfg.jx, cfg.ix))


def write(outdir: str, nout: int, cfg: PIPConfig, U_h, U_m, time: float, ac=None, xi=None,
          gm_ion=None, gm_rec=None, arb_heat=None, **sav) -> str:
    """Write <outdir>/t<nout>.h5; returns its path.  U_h (5, jx, ix) and
    U_m (nvar_m, jx, ix) are the state planes in memory order."""
    U_h = np.ascontiguousarray(U_h, dtype=np.float64)
    U_m = np.ascontiguousarray(U_m, dtype=np.float64)
    if U_h.shape != (5, cfg.jx, cfg.ix) or U_m.shape != (cfg.nvar_m, cfg.jx, cfg.ix):
        raise ValueError("state shapes do not match the configuration")
    grid = grid_arrays(cfg)
    planes = [_plane(cfg, a) for a in (ac, xi, gm_ion, gm_rec, arb_heat)]
    f = PipioFields()
    f.time = float(time)
    for name, a in zip(("x", "y", "z", "dx", "dy", "dz"), grid):
        setattr(f, name, _ptr(a))
    f.U_h, f.U_m = _ptr(U_h), _ptr(U_m)
    for name, a in zip(("ac", "xi", "gm_ion", "gm_rec", "arb_heat"), planes):
        setattr(f, name, _ptr(a))
    L = layout(cfg, **sav)
    _check(lib().pipio_write(outdir.encode(), int(nout), ctypes.byref(L), ctypes.byref(f)))
    return os.path.join(outdir, f"t{int(nout):04d}.h5")


def read(indir: str, nout: int, cfg: PIPConfig, ac: bool = False, rates: bool = False):
    """Restart read of <indir>/t<nout>.h5: dict with time, the grid arrays,
    U_h, U_m and, when asked for, ac and the rate fields (ion, rec, aheat)."""
    out = {k: np.empty(n) for k, n in (("x", cfg.ix), ("y", cfg.jx if cfg.ndim >= 2 else 1),
                                         ("z", 1), ("dx", cfg.ix), ("dy", cfg.jx if cfg.ndim >= 2 else 1),
                                         ("dz", 1))}
    out["U_h"] = np.empty((5, cfg.jx, cfg.ix))
    out["U_m"] = np.empty((cfg.nvar_m, cfg.jx, cfg.ix))
    extra = {"ac": ac, "gm_ion": rates, "gm_rec": rates, "arb_heat": rates and cfg.flag_ir_type == 0}
    for k, want in extra.items():
        out[k] = np.empty((cfg.jx, cfg.ix)) if want else None
    t = ctypes.c_double()
    L = layout(cfg)
    _check(lib().pipio_read(indir.encode(), int(nout), ctypes.byref(L), ctypes.byref(t),
                            *[_ptr(out[k]) for k in ("x", "y", "z", "dx", "dy", "dz", "U_h", "U_m",
                                                      "ac", "gm_ion", "gm_rec", "arb_heat")]))
    out["time"] = t.value
    return {k: v for k, v in out.items() if v is not None}


def output_fields(cfg: PIPConfig, ac, rates=None, xi=None) -> dict:
    """The coefficient planes a two-fluid output hands to ``write``: ac and
    xi always (def_varfiles writes both for flag_pip = 1, IO_rot.f90:234-237),
    and with ionisation/recombination the rates gm_ion, gm_rec and, for
    flag_IR_type = 0, arb_heat (IO_rot.f90:238-241, 254-256, 303-306).
    rates = (gm_rec, gm_ion, arb_heat).  Runs without ambipolar diffusion
    never set xi_n (set_xin runs only for flag_amb >= 1,
    Solver_rot.f90:108-110), so the reference writes an unset allocation
    there; the default here is a plane of zeros."""
    xi = np.zeros((cfg.jx, cfg.ix)) if xi is None else xi
    out = dict(ac=ac, xi=xi)
    if cfg.flag_ir:
        gm_rec, gm_ion, arb = rates
        out.update(gm_ion=gm_ion, gm_rec=gm_rec, arb_heat=arb if cfg.flag_ir_type == 0 else None)
    return out


def write_solver(solver, outdir: str, nout: int, xi=None, **sav) -> str:
    """Output of a PIPSolver's current state (downloaded), time and the
    coefficient planes of ``output_fields``."""
    cfg = solver.cfg
    U_h, U_m = solver.state()
    rates = solver.rates() if cfg.flag_ir else None
    return write(outdir, nout, cfg, U_h, U_m, solver.scalars()["time"],
                 **output_fields(cfg, solver.ac_field(), rates, xi), **sav)


def restart_solver(solver, indir: str, nout: int):
    """restart (IO_rot.f90:405-459) into a PIPSolver: uploads the state (and
    the per-cell ac field when the configuration has one), runs the prologue
    the reference runs after reading (Initial_rot.f90:140-160: ghosts,
    coefficients, first dt) and sets the time; returns the snapshot dict."""
    cfg = solver.cfg
    arb = cfg.flag_ir >= 1 and cfg.flag_ir_type == 0
    snap = read(indir, nout, cfg, ac=bool(cfg.ac_field), rates=arb)
    x, y = grid_arrays(cfg)[:2]
    if not (np.allclose(snap["x"], x, rtol=1e-12, atol=0) and np.allclose(snap["y"], y, rtol=1e-12, atol=0)):
        raise PipioError(f"t{int(nout):04d}.h5 holds another grid than the configuration")
    solver.upload(snap["U_h"], snap["U_m"], snap.get("ac"))
    if arb:  # arb_heat comes from the file, not from the restart state
        solver.set_arb_heat(snap["arb_heat"])
    dt = solver.prologue()
    solver.set_time(snap["time"], dt)
    return snap
