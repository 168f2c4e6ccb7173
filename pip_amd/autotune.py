"""Autotuner for the tile shapes of the stage kernels and the ensemble
workgroup width.  Each candidate is timed on the real problem (device time
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
ENS_TILES = [64, 128, 256, 512]


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


def stage_key(cfg) -> str:
    return (f"stage2:{cfg.ndim}:{cfg.nx}x{cfg.ny}:nv{cfg.nvar_m}:ir{cfg.flag_ir}:ac{int(cfg.ac_field)}"
            f":col{cfg.flag_col}:imp{cfg.flag_pip_imp}")


def tune_stage(make_solver, cfg, steps: int = 20, use_cache: bool = True):
    """make_solver((tx, ty, xcd_remap)) -> PIPSolver ready to step.  Returns
    (best (tx, ty, xcd_remap), timings)."""
    cache = _load()
    key = stage_key(cfg)
    if use_cache and key in cache:
        return tuple(cache[key]["tile"]), cache[key]["ms"]
    timings = {}
    for tile in STAGE_TILES:
        for remap in XCD_REMAP:
            s = make_solver((tile[0], tile[1], remap))
            s.enqueue(3)
            s.sync()
            ev = s.time_steps(steps)
            timings[f"{tile[0]}x{tile[1]}x{remap}"] = ev["total_ms"] / steps
            s.close()
    best = min(timings, key=timings.get)
    tile = tuple(int(v) for v in best.split("x"))
    cache[key] = {"tile": list(tile), "ms": timings}
    _save(cache)
    return tile, timings


def tune_ensemble(ens, key: str, use_cache: bool = True):
    """Time the whole ensemble for each workgroup width; returns (best, timings)."""
    cache = _load()
    if use_cache and key in cache:
        ens.set_tile(cache[key]["tile"])
        return cache[key]["tile"], cache[key]["ms"]
    timings = {}
    for tx in ENS_TILES:
        ens.set_tile(tx)
        ens.reset()
        timings[str(tx)] = ens.run(graph=False)
    best = int(min(timings, key=timings.get))
    ens.set_tile(best)
    cache[key] = {"tile": best, "ms": timings}
    _save(cache)
    return best, timings
