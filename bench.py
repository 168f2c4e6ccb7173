"""Benchmark of the fused two-fluid step on one MI355X.

Workload: the 2-D two-fluid Orszag-Tang problem on a 1024^2 periodic box in
FP64 (xi_n=0.9, alpha_c=1e3, Dedner psi, SSP-RK2, CFL 0.3), 500 steps: one
"step" is one full pass of the time loop (two fused stage kernels + the dt
finalisation) over the whole grid, with the state resident in HBM.

Prints one JSON line (see the repository documentation for the fields):
value = interior cells x steps / wall time of the timed steps;
roofline = algorithmic HBM bytes of the dominant kernel (the stage-2 kernel:
3*nv*8 B per cell) / its average duration measured with hipEvents on the
library's stream over the same timed steps; cpu_baseline = the CPU oracle
(single core) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_PMC = os.path.join(ROOT, "profiles", "stage2_orszag_tang_1024_pmc.md")


def traffic_from_profile():
    """HBM bytes per stage-2 launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    if not os.path.exists(PROFILE_PMC):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(PROFILE_PMC).read())
    return float(m.group(1)) if m else None


def cpu_baseline(n: int, steps: int):
    from oracle.oracle import Oracle
    from pip_amd import problems
    cfg, uh, um, ac = problems.orszag_tang(n)
    o = Oracle(cfg, uh, um, ac)
    t0 = time.perf_counter()
    o.run(nsteps=steps)
    dt = time.perf_counter() - t0
    return {"value": cfg.ncells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle (C restatement of the reference Fortran), Orszag-Tang {n}x{n}, "
                      f"{steps} steps, {dt:.1f} s on one core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="auto", help="TXxTY, or auto (autotuner)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: sync around the timed region
    from pip_amd import autotune, problems
    from pip_amd.solver import PIPSolver

    cfg, uh, um, ac = problems.orszag_tang(args.n)
    if args.tile == "auto":
        tile, _ = autotune.tune_stage(lambda t: PIPSolver(cfg, uh, um, ac, tile=t), cfg)
    else:
        tile = tuple(int(v) for v in args.tile.split("x"))
    s = PIPSolver(cfg, uh, um, ac, tile=tile)
    s.enqueue(args.warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev = s.time_steps(args.steps)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = s.status()
    assert st == 0, f"stop_sim flags {st}"

    ncell = cfg.ncells
    value = ncell * args.steps / wall
    nv = cfg.nvar_m + cfg.nvar_h
    bytes_stage2 = 3 * nv * 8 * ncell
    bytes_stage1 = 2 * nv * 8 * ncell
    t2 = ev["stage2_ms"] / args.steps * 1e-3
    t1 = ev["stage1_ms"] / args.steps * 1e-3
    ach = bytes_stage2 / t2 / 1e9
    out = {
        "metric": "FP64 two-fluid cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Orszag-Tang initial state, no checkpoint)",
        "config": {"workload": "orszag_tang_2d_two_fluid", "nx": args.n, "ny": args.n,
                   "nvar": nv, "xi_n": 0.9, "alpha_c": 1.0e3, "flag_divb": 1,
                   "ssp_rk": 2, "cfl": 0.3, "tile": list(tile)},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic_from_profile(),
                     "kernel": "k_stage (stage 2)", "algorithmic_bytes_per_launch": bytes_stage2,
                     "avg_launch_ms": t2 * 1e3,
                     "stage1_avg_launch_ms": t1 * 1e3,
                     "stage1_achieved_GBs": bytes_stage1 / t1 / 1e9},
        "hbm_roofline_cell_updates_frac": value * (5 * nv * 8) / (HBM_PEAK_GBS * 1e9),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
