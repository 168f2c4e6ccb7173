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
Two more objects compare the training-system patterns with plain runs:
batched = 256 uneven 1-D shock tubes in one persistent hipGraph launch
against the same 256 runs as sequential plain pip_advance calls (results
compared bit for bit); adjoint = gradient (checkpointed forward + reverse
sweep) of the 512^2 Alfven-wave damping run (200 steps, checkpoints every
20) against the plain forward run.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

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


def batched_leg(n_members: int = 256, tend: float = 0.1, max_steps: int = 20000):
    """Ensemble of n_members 1-D Brio-Wu-type shock tubes (nx log-uniform in
    [256, 4096], xi_n, alpha_c and the pressure ratio drawn from seed 7) run
    to t = tend: one persistent launch vs sequential plain launches."""
    import torch
    from pip_amd import autotune, problems
    from pip_amd.ensemble import Ensemble
    from pip_amd.solver import PIPSolver
    params = problems.ensemble_members(n_members, seed=7)
    mem = [problems.ensemble_member(p)[:3] for p in params]
    ens = Ensemble(mem, tend=tend, max_steps=max_steps)
    tile, tune_ms = autotune.tune_ensemble(ens, f"ens:{n_members}:seed7:tend{tend}")
    ens.reset()
    ens.run(graph=True)  # builds and caches the graph
    ens.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens_ms = ens.run(graph=True)
    wall_ens = time.perf_counter() - t0
    updates = 0
    wall_seq = 0.0
    identical = True
    for k, (cfg, uh, um) in enumerate(mem):
        s = PIPSolver(cfg, uh, um)
        s.sync()
        t0 = time.perf_counter()
        n, t = s.advance(nsteps=max_steps, tend=tend)
        wall_seq += time.perf_counter() - t0
        r = ens.result(k)
        updates += cfg.nx * n
        eh, em = ens.state(k)
        ph, pm = s.state()
        identical &= (r["nt"] == n and r["time"] == t and np.array_equal(em[:, :, 2:-2], pm[:, :, 2:-2])
                      and np.array_equal(eh[:, :, 2:-2], ph[:, :, 2:-2]))
        s.close()
    ens.close()
    return {"workload": f"{n_members} x 1-D two-fluid shock tube, nx log-uniform [256,4096], t_end={tend}",
            "cell_updates": updates, "tile": tile, "tile_timings_ms": tune_ms,
            "ensemble_s": wall_ens, "ensemble_device_ms": ens_ms, "sequential_plain_s": wall_seq,
            "ensemble_value": updates / wall_ens, "sequential_value": updates / wall_seq,
            "unit": "cell-updates/s", "speedup": wall_seq / wall_ens, "bitwise_identical": bool(identical)}


def adjoint_leg(n: int = 512, nsteps: int = 200, ckpt: int = 20):
    """Gradient of J = |U_T|^2/2 (interior) w.r.t. the initial state and the
    per-cell alpha_c field vs the plain forward run (pip_advance)."""
    import torch
    from pip_amd import problems
    from pip_amd.adjoint import TwoFluidRun
    from pip_amd.solver import PIPSolver
    cfg, uh, um, ac = problems.alfven_damping(n)
    s = PIPSolver(cfg, uh, um, ac)
    s.advance(5)  # warm-up
    s.close()
    s = PIPSolver(cfg, uh, um, ac)
    s.sync()
    t0 = time.perf_counter()
    s.advance(nsteps)
    t_fwd = time.perf_counter() - t0
    s.close()
    dev = torch.device("cuda:0")
    run = TwoFluidRun(cfg, nsteps, ckpt_every=ckpt)
    U0 = torch.tensor(np.concatenate([um, uh]), device=dev)
    acT = torch.tensor(ac, device=dev)
    mask = torch.zeros((cfg.jx, cfg.ix), dtype=torch.float64, device=dev)
    mask[2:-2, 2:-2] = 1.0
    times = []
    for _ in range(2):  # the first pass allocates the adjoint buffers
        U = U0.clone().requires_grad_(True)
        a = acT.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        UT = run.run(U, a)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        UT.backward(UT.detach() * mask)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t1))
    t_rec, t_bwd = times[-1]
    run.close()
    return {"workload": f"Alfven-wave damping {n}x{n}, {nsteps} SSP-RK2 steps, checkpoints every {ckpt}",
            "forward_s": t_fwd, "forward_recorded_s": t_rec, "backward_s": t_bwd,
            "gradient_over_forward": (t_rec + t_bwd) / t_fwd, "backward_over_forward": t_bwd / t_fwd}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="auto", help="TXxTY, or auto (autotuner)")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and adjoint legs")
    ap.add_argument("--only", choices=["batched", "adjoint"], help="run one extra leg alone (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: sync around the timed region
    from pip_amd import autotune, problems
    from pip_amd.solver import PIPSolver

    if args.only:
        print(json.dumps({args.only: batched_leg() if args.only == "batched" else adjoint_leg()}))
        return

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
    if not args.no_extra:
        out["batched"] = batched_leg()
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
