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
# committed rocprofv3 --pmc summaries of the stage-2 kernel: the one-pass
# k_stage and the fluid-by-fluid k_stage2p
PROFILE_PMC = {False: os.path.join(ROOT, "profiles", "stage2_orszag_tang_1024_pmc.md"),
               True: os.path.join(ROOT, "profiles", "stage2p_orszag_tang_1024_pmc.md")}


def two_pass_stage2(tile) -> bool:
    return len(tile) > 3 and bool((int(tile[3]) >> 1) & 1)


def traffic_from_profile(tile):
    """(HBM bytes per stage-2 launch, tile it was measured with) from the
    committed summary of the stage-2 kernel the tile selects (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = PROFILE_PMC[two_pass_stage2(tile)]
    if not os.path.exists(path):
        return None, None
    txt = open(path).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    t = re.search(r"measured with tile:\s*(\S+)", txt)
    return (float(m.group(1)) if m else None), (t.group(1) if t else None)


def _pmc_rows(d: str, counter: str):
    """(kernel name, value) per dispatch of one counter from the rocprofv3
    --pmc CSVs under d, in dispatch order."""
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") == counter:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def write_pmc_summary(d: str, cmds: dict, tile: str):
    """HBM bytes per stage-2 k_stage launch from four rocprofv3 --pmc runs
    (d/fetch, d/write: the main workload; d/copy_fetch, d/copy_write: the
    1 GiB copy).  FETCH_SIZE / WRITE_SIZE are in KiB per dispatch; each is
    scaled by known bytes / counted bytes of the copy kernel."""
    def stage2(counter, sub):
        v = [val for k, val in _pmc_rows(os.path.join(d, sub), counter) if "k_stage" in k]
        return v[1::2], v[0::2]  # each step launches stage 1 then stage 2

    def copy_kib(counter, sub):
        v = [val for k, val in _pmc_rows(os.path.join(d, sub), counter) if "k_stage" not in k]
        return max(v) if v else float("nan")
    gib_kib = float(1 << 20)
    f_scale = gib_kib / copy_kib("FETCH_SIZE", "copy_fetch")
    w_scale = gib_kib / copy_kib("WRITE_SIZE", "copy_write")
    f2, f1 = stage2("FETCH_SIZE", "fetch")
    w2, w1 = stage2("WRITE_SIZE", "write")
    mean = lambda x: sum(x) / len(x)
    fetch_b = mean(f2) * 1024 * f_scale
    write_b = mean(w2) * 1024 * w_scale
    ncell, nv = 1024 * 1024, 14
    tt = tuple(int(v) for v in tile.split("x"))
    kern = "k_stage2p (fluid by fluid)" if two_pass_stage2(tt) else "k_stage (one pass)"
    lines = [f"# Stage-2 kernel {kern}, Orszag-Tang 1024^2 FP64: HBM traffic per launch", "",
             f"measured with tile: {tile}", "",
             "(each step launches stage 1 then stage 2)", "",
             "Commands (each its own pass):", ""]
    lines += [f"    {c}" for c in cmds.values()]
    lines += ["", f"copy calibration: FETCH_SIZE scale {f_scale:.4f}, WRITE_SIZE scale {w_scale:.4f} "
              "(1 GiB read + 1 GiB written per copy)", "",
              f"stage-2 launches: {len(f2)}; mean FETCH_SIZE {mean(f2):.1f} KiB, WRITE_SIZE {mean(w2):.1f} KiB",
              f"stage-1 launches: {len(f1)}; mean FETCH_SIZE {mean(f1):.1f} KiB, WRITE_SIZE {mean(w1):.1f} KiB", "",
              f"stage-2 HBM bytes per launch (scaled): read {fetch_b:.4e}, write {write_b:.4e}",
              f"algorithmic bytes per launch: read {2 * nv * 8 * ncell:.4e}, write {nv * 8 * ncell:.4e}", "",
              f"traffic_bytes_per_launch: {fetch_b + write_b:.6e}", ""]
    open(PROFILE_PMC[two_pass_stage2(tt)], "w").write("\n".join(lines))
    print("\n".join(lines))


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
    to t = tend: the batched ensemble (launch mode and width autotuned) vs
    sequential plain launches."""
    import torch
    from pip_amd import autotune, problems
    from pip_amd.ensemble import Ensemble
    from pip_amd.solver import PIPSolver
    params = problems.ensemble_members(n_members, seed=7)
    mem = [problems.ensemble_member(p)[:3] for p in params]
    ens = Ensemble(mem, tend=tend, max_steps=max_steps)
    (mode, tile), tune_ms = autotune.tune_ensemble(ens, f"ens:{n_members}:seed7:tend{tend}")
    ens.reset()
    ens.run(graph=True)  # builds and caches the graph (persistent mode)
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
            "cell_updates": updates, "mode": mode, "tile": tile, "tile_timings_ms": tune_ms,
            "ensemble_s": wall_ens, "ensemble_device_ms": ens_ms, "sequential_plain_s": wall_seq,
            "ensemble_value": updates / wall_ens, "sequential_value": updates / wall_seq,
            "unit": "cell-updates/s", "speedup": wall_seq / wall_ens, "bitwise_identical": bool(identical)}


def ensemble_leg(mode: str, tile: int, n_members: int = 256, tend: float = 0.1, max_steps: int = 20000):
    """The batched leg's ensemble alone in one launch mode and width, run
    twice from its initial state (for profiles)."""
    from pip_amd import problems
    from pip_amd.ensemble import Ensemble
    params = problems.ensemble_members(n_members, seed=7)
    mem = [problems.ensemble_member(p)[:3] for p in params]
    ens = Ensemble(mem, tend=tend, max_steps=max_steps, tile=tile)
    ens.set_mode(mode)
    runs = []
    for _ in range(2):
        ens.reset()
        t0 = time.perf_counter()
        ms = ens.run(graph=True)
        runs.append({"device_ms": ms, "wall_s": time.perf_counter() - t0})
    steps = max(ens.result(k)["nt"] for k in range(n_members))
    ens.close()
    return {"workload": f"{n_members} x 1-D two-fluid shock tube, seed 7, t_end={tend}", "mode": mode,
            "tile": tile, "longest_run_steps": steps, "runs": runs}


def adjoint_leg(n: int = 512, nsteps: int = 200, ckpt: int = 20):
    """Gradient of J = |U_T|^2/2 (interior) w.r.t. the initial state and the
    per-cell alpha_c field vs the plain forward run (pip_advance)."""
    import torch
    from pip_amd import autotune, problems
    from pip_amd.adjoint import TwoFluidRun
    from pip_amd.solver import PIPSolver
    cfg, uh, um, ac = problems.alfven_damping(n)
    # the forward and the gradient run the same (autotuned) stage kernels
    tile, _ = autotune.tune_stage(lambda t: PIPSolver(cfg, uh, um, ac, tile=t), cfg)
    s = PIPSolver(cfg, uh, um, ac, tile=tile)
    s.advance(5)  # warm-up
    s.close()
    s = PIPSolver(cfg, uh, um, ac, tile=tile)
    s.sync()
    t0 = time.perf_counter()
    s.advance(nsteps)
    t_fwd = time.perf_counter() - t0
    s.close()
    dev = torch.device("cuda:0")
    run = TwoFluidRun(cfg, nsteps, ckpt_every=ckpt, tile=tile)
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
            "tile": list(tile),
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
    ap.add_argument("--tile", type=str, default="auto", help="TXxTY[xXCD_ORDER[xTWO_PASS_STAGE_MASK]], or auto (autotuner)")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and adjoint legs")
    ap.add_argument("--pmc-summary", metavar="DIR", help="write the HBM-traffic profile summary from "
                    "rocprofv3 --pmc runs under DIR (fetch, write, copy_fetch, copy_write; the main "
                    "runs used --tile) and exit")
    ap.add_argument("--only", choices=["batched", "adjoint", "copy", "ensemble"],
                    help="run one extra leg alone (profiling); copy = 1 GiB device copy that "
                         "calibrates the HBM byte counters; ensemble = the batched leg's ensemble "
                         "alone, twice, in --ens-mode at --ens-tile (no autotuning, no plain runs)")
    ap.add_argument("--ens-mode", choices=["persistent", "lockstep"], default="lockstep")
    ap.add_argument("--ens-tile", type=int, choices=[64, 128, 256, 512], default=256)
    args = ap.parse_args()
    if args.pmc_summary:
        pm = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 60 rocprofv3 --pmc {c} --output-format csv "
              "-d {d}/{sub} -o run -- python3 /root/repo/bench.py {a}")
        tile = "32x7" if args.tile == "auto" else args.tile
        main_args = f"--steps 20 --warmup 5 --no-cpu --no-extra --tile {tile}"
        cmds = {sub: pm.format(c=c, d=args.pmc_summary, sub=sub, a=a) for sub, c, a in (
            ("fetch", "FETCH_SIZE", main_args), ("write", "WRITE_SIZE", main_args),
            ("copy_fetch", "FETCH_SIZE", "--only copy"), ("copy_write", "WRITE_SIZE", "--only copy"))}
        write_pmc_summary(args.pmc_summary, cmds, tile)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: sync around the timed region
    from pip_amd import autotune, problems
    from pip_amd.solver import PIPSolver

    if args.only == "copy":
        n = 1 << 27  # 2^27 doubles = 1 GiB per buffer
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy": {"bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n}}))
        return
    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_leg(args.ens_mode, args.ens_tile)}))
        return
    if args.only:
        print(json.dumps({args.only: batched_leg() if args.only == "batched" else adjoint_leg()}))
        return

    cfg, uh, um, ac = problems.orszag_tang(args.n)
    if args.tile == "auto":
        tile, _ = autotune.tune_stage(lambda t: PIPSolver(cfg, uh, um, ac, tile=t), cfg)
    else:
        tile = tuple(int(v) for v in args.tile.split("x"))  # TXxTY[xREMAP[xTWO_PASS_MASK]]
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

    # measured copy bandwidth on this GPU (1 GiB read + 1 GiB written)
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * 8 * (1 << 27) / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst

    ncell = cfg.ncells
    value = ncell * args.steps / wall
    nv = cfg.nvar_m + cfg.nvar_h
    bytes_stage2 = 3 * nv * 8 * ncell
    bytes_stage1 = 2 * nv * 8 * ncell
    t2 = ev["stage2_ms"] / args.steps * 1e-3
    t1 = ev["stage1_ms"] / args.steps * 1e-3
    ach = bytes_stage2 / t2 / 1e9
    traffic = traffic_from_profile(tile)
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
                   "ssp_rk": 2, "cfl": 0.3, "tile": list(tile[:2]),
                   "xcd_remap": int(tile[2]) if len(tile) > 2 else 0,
                   "two_pass_stages": int(tile[3]) if len(tile) > 3 else 0},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic[0], "traffic_tile": traffic[1],
                     "kernel": ("k_stage2p" if len(tile) > 3 and (tile[3] >> 1) & 1 else "k_stage") + " (stage 2)",
                     "algorithmic_bytes_per_launch": bytes_stage2,
                     "avg_launch_ms": t2 * 1e3,
                     "stage1_avg_launch_ms": t1 * 1e3,
                     "stage1_achieved_GBs": bytes_stage1 / t1 / 1e9,
                     "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": ach / copy_gbs},
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
