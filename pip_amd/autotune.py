"""Autotuner for the tile shapes of the stage kernels and the ensemble
launch mode and workgroup width.  Each candidate is timed on the real problem (device time
from hipEvents) and the fastest is cached in pip_amd/lib/autotune.json,
keyed by problem signature; results are identical for every candidate, so
the choice affects speed only."""
from __future__ import annotations

import json
import os

from . import build

CACHE = os.path.join(build.LIBDIR, "autotune.json")
STAGE_TILES = [(16, 15), (32, 7), (32, 15), (64, 7), (32, 8), (16, 7)]
XCD_REMAP = [0, 1]  # row-major tile order, or contiguous tile ranges per XCD
TWO_PASS_TILES = [(16, 15), (32, 7), (16, 7)]  # tiles of the fluid-by-fluid stage kernel
RELAXED_TILES = [(16, 15), (32, 7), (16, 7)]  # tiles of the relaxed-arithmetic stage kernel
ENS_TILES = [64, 128, 256, 512]
# one launch; every member spread over workgroups in host-driven chunks; the
# same chunks captured as a hipGraph
ENS_MODES = ["persistent", "lockstep", "graph"]


def _load():
    if os.path.exists(CACHE):
        try:
            return json.load(open(CACHE))
        except Exception:
            return {}
    return {}


def _save(d):
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    json.dump(d, open(CACHE, "w"), indent=1)


def stage_key(cfg, arith: str = "exact") -> str:
    key = (f"stage2:{cfg.ndim}:{cfg.nx}x{cfg.ny}:nv{cfg.nvar_m}:ir{cfg.flag_ir}:ac{int(cfg.ac_field)}"
           f":col{cfg.flag_col}:imp{cfg.flag_pip_imp}:twopass")
    return key if arith == "exact" else key + ":" + arith


def two_pass_ok(cfg) -> bool:
    """Configurations the fluid-by-fluid stage kernel covers."""
    return cfg.ndim == 2 and cfg.flag_col == 1 and cfg.flag_pip_imp == 0


def tune_stage(make_solver, cfg, steps: int = 20, use_cache: bool = True, arith: str = "exact"):
    """make_solver((tx, ty, xcd_remap, two_pass)) -> PIPSolver ready to step
    (two_pass: mask of the SSP-RK stages that take the fluids one at a
    time) in the given arithmetic (the relaxed kernels come in three tiles,
    one pass).  Returns (best (tx, ty, xcd_remap, two_pass), timings)."""
    cache = _load()
    key = stage_key(cfg, arith)
    if use_cache and key in cache:
        return tuple(cache[key]["tile"]), cache[key]["ms"]
    timings = {}
    if arith == "relaxed":
        # 1-D runs have one kernel (255-cell tiles); the tile setting is unused there
        cands = [(t[0], t[1], r, 0) for t in RELAXED_TILES for r in XCD_REMAP] if cfg.ndim == 2 else [(16, 15, 0, 0)]
    else:
        cands = [(t[0], t[1], r, 0) for t in STAGE_TILES for r in XCD_REMAP]
    if arith == "exact" and two_pass_ok(cfg):
        full = (1 << cfg.t_order) - 1
        cands += [(t[0], t[1], r, m) for t in TWO_PASS_TILES for r in XCD_REMAP
                  for m in sorted({1 << (cfg.t_order - 1), full})]
    for cand in cands:
        s = make_solver(cand)
        s.enqueue(3)
        s.sync()
        ev = s.time_steps(steps)
        timings["x".join(str(v) for v in cand)] = ev["total_ms"] / steps
        s.close()
    best = min(timings, key=timings.get)
    tile = tuple(int(v) for v in best.split("x"))
    cache[key] = {"tile": list(tile), "ms": timings}
    _save(cache)
    return tile, timings


def tune_ensemble(ens, key: str, use_cache: bool = True):
    """Time the whole ensemble for each launch mode and workgroup width;
    sets the fastest and returns ((mode, width), timings)."""
    cache = _load()
    key = key + ":modes"
    if use_cache and key in cache:
        mode, tx = cache[key]["best"]
        ens.set_mode(mode)
        ens.set_tile(tx)
        return (mode, tx), cache[key]["ms"]
    timings = {}
    for mode in ENS_MODES:
        ens.set_mode(mode)
        for tx in ENS_TILES:
            ens.set_tile(tx)
            if mode == "graph":  # the first run captures the graph
                ens.reset()
                ens.run(graph=False)
            ens.reset()
            timings[f"{mode}:{tx}"] = ens.run(graph=False)
    best = min(timings, key=timings.get)
    mode, tx = best.split(":")[0], int(best.split(":")[1])
    ens.set_mode(mode)
    ens.set_tile(tx)
    cache[key] = {"best": [mode, tx], "ms": timings}
    _save(cache)
    return (mode, tx), timings
