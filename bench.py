"""Benchmark of the fused two-fluid step on one MI355X.

Headline workload (default): the 2-D two-fluid Kelvin-Helmholtz problem with
temperature-dependent ionisation/recombination (set_IR case 3, ion_pot /
arb_heat energy terms) at 2048^2 in FP64: x periodic, y reflecting, xi_n=0.8,
alpha_c=1e2, Dedner psi, SSP-RK2, CFL 0.3, 1000 steps.  One "step" is one
full pass of the time loop (two fused stage kernels + the dt finalisation)
over the whole grid, with the state resident in HBM.

Prints one JSON line:
  value      interior cells x steps / wall time of the timed steps;
  roofline   the stage kernel k_stage (both SSP-RK stages are launches of
             it): algorithmic HBM bytes of a step (stage 1 reads U^n and
             writes U^(1); stage 2 reads U^(1), U^n and writes U^(n+1);
             every stage reads the arb_heat field) over the summed duration
             of its two launches, both measured with hipEvents on the
             library's stream over the timed steps; traffic = measured HBM
             bytes per launch from the committed rocprofv3 summary;
  cpu_baseline   the CPU oracle (the C restatement of the reference path) on
             one host core, median of three timed steps of the same
             workload; cpu_baseline_threads the same on all host cores
             (OpenMP, up to 16 threads);
  secondary  the Orszag-Tang 1024^2 run (the 2-D configuration without IR);
  batched    256 uneven 1-D shock tubes as one graph-captured ensemble
             against the same runs as sequential plain launches;
  adjoint    gradient (checkpointed forward + reverse sweep) of the 512^2
             Alfven-wave damping run (200 steps, checkpoints every 20)
             against the plain forward run.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: (problem builder name, default n, config fields of the JSON line)
    "kh": ("kelvin_helmholtz", 2048),
    "ot": ("orszag_tang", 1024),
}
WL_NAME = {"kh": "kelvin_helmholtz_2d_ir", "ot": "orszag_tang_2d_two_fluid"}


def pmc_profile_path(wl: str, n: int, arith: str = "exact") -> str:
    tag = "" if arith == "exact" else f"_{arith}"
    return os.path.join(ROOT, "profiles", f"k_stage_{WL_NAME[wl]}_{n}{tag}_pmc.md")


# the tests that hold each arithmetic to the oracle on the headline workload
PARITY_TEST = {
    "exact": "tests/test_gpu_parity_full.py::test_kelvin_helmholtz_ir_2048_100_steps (bitwise, 100 steps at "
             "2048^2), ::test_kelvin_helmholtz_ir_384_1000_steps (bitwise, 1000 steps at 384^2) and "
             "tests/test_gpu_parity_long.py (opt-in: bitwise, all 1000 steps at 2048^2)",
    "relaxed": "tests/test_gpu_parity_full.py::test_kelvin_helmholtz_ir_2048_100_steps (after 100 steps "
               "relative L1 <= 1e-10 on the order-one variables, <= 1e-5 on the perturbation-born ones; "
               "does not meet a per-variable 1e-12 bar) and tests/test_gpu_relaxed.py (shock positions)",
}


def make_problem(wl: str, n: int):
    from pip_amd import problems
    return getattr(problems, WORKLOADS[wl][0])(n)


def workload_config(wl: str, n: int, cfg, tile):
    c = {"workload": WL_NAME[wl], "nx": n, "ny": n, "nvar": cfg.nvar_m + cfg.nvar_h,
         "flag_divb": cfg.flag_divb, "ssp_rk": cfg.t_order, "cfl": cfg.safety,
         "alpha_c": cfg.col, "flag_bnd": list(cfg.flag_bnd)}
    if wl == "kh":
        c.update({"xi_n": 0.8, "flag_ir": cfg.flag_ir, "flag_ir_type": cfg.flag_ir_type,
                  "shear": "vx = 0.5 tanh(y/0.05)", "bx": 0.1, "seed": 20241202})
    else:
        c.update({"xi_n": 0.9})
    c.update({"tile": list(tile[:2]), "xcd_remap": int(tile[2]) if len(tile) > 2 else 0,
              "two_pass_stages": int(tile[3]) if len(tile) > 3 else 0})
    return c


def traffic_from_profile(wl: str, n: int, arith: str = "exact"):
    """(HBM bytes per k_stage launch averaged over both stages, per stage-1
    and per stage-2 launch, tile measured) from the committed summary."""
    path = pmc_profile_path(wl, n, arith)
    if not os.path.exists(path):
        return None
    txt = open(path).read()

    def num(key):
        m = re.search(key + r":\s*([0-9.eE+]+)", txt)
        return float(m.group(1)) if m else None
    t = re.search(r"measured with tile:\s*(\S+)", txt)
    return {"mean": num("traffic_bytes_per_launch"), "stage1": num("stage1_bytes_per_launch"),
            "stage2": num("stage2_bytes_per_launch"), "tile": t.group(1) if t else None}


def valu_from_profile(wl: str, n: int, arith: str = "exact"):
    """(VALU lane-instructions per cell-update, the FP64 VALU ceiling in
    cell-updates/s they imply) from the committed SQ counter summary."""
    tag = "" if arith == "exact" else f"_{arith}"
    path = os.path.join(ROOT, "profiles", f"k_stage_{WL_NAME[wl]}_{n}{tag}_sq.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    li = re.search(r"lane-instructions per cell-update \(both stages\):\s*([0-9.eE+]+)", txt)
    ce = re.search(r"FP64 VALU ceiling .*?:\s*([0-9.eE+]+) cell-updates/s", txt)
    if not (li and ce):
        return None
    return float(li.group(1)), float(ce.group(1))


def _csv_rows(d: str, pattern: str):
    import csv
    import glob
    out = []
    for f in sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True)):
        out += list(csv.DictReader(open(f)))
    return out


def _pmc_rows(d: str, counter: str):
    """(kernel name, value) per dispatch of one counter from the rocprofv3
    --pmc CSVs under d, in dispatch order."""
    rows = [(int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"]))
            for r in _csv_rows(d, "*counter_collection.csv") if r.get("Counter_Name") == counter]
    rows.sort()
    return [(k, v) for _, k, v in rows]


def resolved_tile(wl: str, n: int, tile: str, arith: str = "exact") -> str:
    """The tile string a run with --tile `tile` uses (auto: the autotuner's
    cached choice for the workload, if any)."""
    if tile != "auto":
        return tile
    from pip_amd import autotune
    cfg = make_problem(wl, min(n, 64))[0].replace(nx=n, ny=n)
    hit = autotune._load().get(autotune.stage_key(cfg, arith))
    return "x".join(str(v) for v in hit["tile"]) if hit else "auto (no cached choice)"


def write_pmc_summary(d: str, cmds: dict, wl: str, n: int, tile: str, arith: str = "exact"):
    """HBM bytes per k_stage launch from four rocprofv3 --pmc runs (d/fetch,
    d/write: the workload; d/copy_fetch, d/copy_write: the 1 GiB copy).
    FETCH_SIZE / WRITE_SIZE are in KiB per dispatch; each is scaled by known
    bytes / counted bytes of the copy kernel."""
    def stages(counter, sub):
        v = [val for k, val in _pmc_rows(os.path.join(d, sub), counter) if "k_stage" in k]
        return v[0::2], v[1::2]  # each step launches stage 1 then stage 2

    def copy_kib(counter, sub):
        v = [val for k, val in _pmc_rows(os.path.join(d, sub), counter) if "k_stage" not in k]
        return max(v) if v else float("nan")
    gib_kib = float(1 << 20)
    f_scale = gib_kib / copy_kib("FETCH_SIZE", "copy_fetch")
    w_scale = gib_kib / copy_kib("WRITE_SIZE", "copy_write")
    f1, f2 = stages("FETCH_SIZE", "fetch")
    w1, w2 = stages("WRITE_SIZE", "write")
    mean = lambda x: sum(x) / len(x)
    b1 = (mean(f1) * f_scale + mean(w1) * w_scale) * 1024
    b2 = (mean(f2) * f_scale + mean(w2) * w_scale) * 1024
    ncell = n * n
    nv = 14  # both workloads: 9 plasma planes (with psi) + 5 neutral planes
    extra = 8 if wl == "kh" else 0  # arb_heat read per stage
    alg1 = (2 * nv * 8 + extra) * ncell
    alg2 = (3 * nv * 8 + extra) * ncell
    lines = [f"# k_stage ({arith} arithmetic), {WL_NAME[wl]} {n}^2 FP64: HBM traffic per launch", "",
             f"measured with tile: {resolved_tile(wl, n, tile, arith)}", "",
             "(each step launches stage 1 then stage 2 of the same kernel)", "",
             "Commands (each its own pass, from the repository root):", ""]
    lines += [f"    {c}" for c in cmds.values()]
    lines += ["", f"copy calibration: FETCH_SIZE scale {f_scale:.4f}, WRITE_SIZE scale {w_scale:.4f} "
              "(1 GiB read + 1 GiB written per copy)", "",
              f"stage-1 launches: {len(f1)}; mean FETCH_SIZE {mean(f1):.1f} KiB, WRITE_SIZE {mean(w1):.1f} KiB",
              f"stage-2 launches: {len(f2)}; mean FETCH_SIZE {mean(f2):.1f} KiB, WRITE_SIZE {mean(w2):.1f} KiB", "",
              f"stage1_bytes_per_launch: {b1:.6e}   (algorithmic {alg1:.6e}, ratio {b1 / alg1:.3f})",
              f"stage2_bytes_per_launch: {b2:.6e}   (algorithmic {alg2:.6e}, ratio {b2 / alg2:.3f})",
              f"traffic_bytes_per_launch: {(b1 + b2) / 2:.6e}   (mean of the two stages; algorithmic "
              f"{(alg1 + alg2) / 2:.6e})", ""]
    path = pmc_profile_path(wl, n, arith)
    open(path, "w").write("\n".join(lines))
    print("\n".join(lines))


SQ_PASSES = (("sq1", "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS"),
             ("sq2", "SQ_WAVE_CYCLES SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_WAIT_INST_ANY"))


def write_sq_summary(d: str, wl: str, n: int, tile: str, out: str, arith: str = "exact"):
    """Instruction and wave-state counters of k_stage per launch (stage 1
    and stage 2 separately) from the rocprofv3 --pmc passes d/sq1, d/sq2,
    with the FP64 VALU ceiling they imply."""
    main_args = f"--workload {wl} --n {n} --steps 10 --warmup 3 --no-cpu --no-extra --tile {tile} --arith {arith}"
    cmds = [f"cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc {c} --output-format csv "
            f"-d {d}/{sub} -o run -- python3 /root/repo/bench.py {main_args}"
            for sub, c in SQ_PASSES]
    vals = {}
    for sub, cs in SQ_PASSES:
        for c in cs.split():
            v = [val for k, val in _pmc_rows(os.path.join(d, sub), c) if "k_stage" in k]
            vals[c] = (sum(v[0::2]) / len(v[0::2]), sum(v[1::2]) / len(v[1::2]))
    ncell = n * n
    lanes = 256 * 4 * 16 * 2.4e9  # FP64 VALU lane-instructions per second (CU x SIMD x lanes x clock)
    lines = [f"# k_stage ({arith} arithmetic), {WL_NAME[wl]} {n}^2 FP64: instruction and wave-state counters", "",
             f"measured with tile: {resolved_tile(wl, n, tile, arith)}", "",
             "Commands (each its own pass, from the repository root):", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "Per launch (mean over the profiled launches; each step launches stage 1 then stage 2):", "",
              "| counter | stage 1 | stage 2 |", "|---|---|---|"]
    for c, (a, b) in vals.items():
        lines.append(f"| {c} | {a:.4e} | {b:.4e} |")
    w1, w2 = vals["SQ_WAVES"]
    i1, i2 = vals["SQ_INSTS_VALU"]
    lane_per_cell = (i1 + i2) * 64 / ncell
    lines += ["", "Derived:", "",
              f"- VALU instructions per wave: stage 1 {i1 / w1:.0f}, stage 2 {i2 / w2:.0f}",
              f"- VALU lane-instructions per cell-update (both stages): {lane_per_cell:.0f}",
              f"- FP64 VALU ceiling (256 CU x 4 SIMD x 16 lanes x 2.4 GHz = {lanes:.3e} lane-instr/s): "
              f"{lanes / lane_per_cell:.3e} cell-updates/s",
              "- share of a wave's cycles issuing VALU (SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES): "
              f"stage 1 {vals['SQ_ACTIVE_INST_VALU'][0] / vals['SQ_WAVE_CYCLES'][0]:.2f}, "
              f"stage 2 {vals['SQ_ACTIVE_INST_VALU'][1] / vals['SQ_WAVE_CYCLES'][1]:.2f}",
              "- share waiting on a dependency (SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES): "
              f"stage 1 {vals['SQ_WAIT_INST_ANY'][0] / vals['SQ_WAVE_CYCLES'][0]:.2f}, "
              f"stage 2 {vals['SQ_WAIT_INST_ANY'][1] / vals['SQ_WAVE_CYCLES'][1]:.2f}",
              "- lane utilisation (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)): "
              f"stage 1 {vals['SQ_THREAD_CYCLES_VALU'][0] / (64 * vals['SQ_ACTIVE_INST_VALU'][0]):.3f}, "
              f"stage 2 {vals['SQ_THREAD_CYCLES_VALU'][1] / (64 * vals['SQ_ACTIVE_INST_VALU'][1]):.3f}", ""]
    open(out, "w").write("\n".join(lines))
    print("\n".join(lines))


ADJ_SQ_PASSES = (("sq1", "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR"),
                 ("sq2", "SQ_WAVE_CYCLES SQ_ACTIVE_INST_VALU SQ_WAIT_INST_ANY SQ_BUSY_CYCLES"),
                 ("sq3", "SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_TRANS_F64"))
ADJ_KERNELS = (("plasma face VJPs", "k_adj_plasma_faces"), ("gather", "k_adj_gather"),
               ("replay stage", ", 1>(pip::StageArgs)"))


def write_adj_sq_summary(d: str, out: str):
    """Instruction and wave-state counters of the adjoint kernels per launch
    (the plasma face launch, the gather, a replay stage) from the rocprofv3
    --pmc passes d/sq1..sq3 of `bench.py --only adjoint`."""
    cmds = [f"cd /tmp && export TMPDIR=/tmp && timeout -s KILL 300 rocprofv3 --pmc {c} --output-format csv "
            f"-d {d}/{sub} -o run -- python3 /root/repo/bench.py --only adjoint" for sub, c in ADJ_SQ_PASSES]
    vals = {}
    for sub, cs in ADJ_SQ_PASSES:
        for c in cs.split():
            rows = _pmc_rows(os.path.join(d, sub), c)
            vals[c] = [sum(v for k, v in rows if key in k) / max(1, sum(1 for k, _ in rows if key in k))
                       for _, key in ADJ_KERNELS]
    lines = ["# Adjoint kernels, Alfven-wave damping 512^2 FP64: instruction and wave-state counters", "",
             "Commands (each its own pass, from the repository root):", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "Per launch (mean over the profiled launches; SQ_*_CYCLES count 4-cycle units, summed "
              "over the waves):", "", "| counter | " + " | ".join(n for n, _ in ADJ_KERNELS) + " |",
              "|---|" + "---|" * len(ADJ_KERNELS)]
    for c, v in vals.items():
        lines.append(f"| {c} | " + " | ".join(f"{x:.4e}" for x in v) + " |")
    lines += ["", "Derived:", ""]
    for i, (name, _) in enumerate(ADJ_KERNELS):
        w, valu = vals["SQ_WAVES"][i], vals["SQ_INSTS_VALU"][i]
        f64 = sum(vals[c][i] for c in ADJ_SQ_PASSES[2][1].split())
        wc = vals["SQ_WAVE_CYCLES"][i]
        lines.append(f"- {name}: {valu / w:.0f} VALU instructions per wave, {f64 / valu:.2f} of them FP64; "
                     f"a wave issues VALU in {vals['SQ_ACTIVE_INST_VALU'][i] / wc:.2f} of its cycles and waits "
                     f"on memory or LDS (s_waitcnt) in {vals['SQ_WAIT_INST_ANY'][i] / wc:.2f}")
    lines.append("")
    open(out, "w").write("\n".join(lines))
    print("\n".join(lines))


# The adjoint leg's checkpoint schedules, in the order adjoint_leg times
# them (4 gradients each): (key, label, keep_segments, overlap_replay);
# keep_segments None is TwoFluidRun's default (2 overlapped, 1 in line).
ADJ_SCHEDULES = [("default", "2 newest segments kept, replays beside the adjoint steps (default)", None, True),
                 ("keep_0", "no segment kept, replays beside the adjoint steps", 0, True),
                 ("keep_5", "5 segments kept, replays beside the adjoint steps", 5, True),
                 ("in_line", "newest segment kept, replays in line", None, False)]
ADJ_GRADIENTS = 4


def _adj_kept(keep, overlap, nseg):
    from pip_amd.adjoint import default_keep_segments
    return min(nseg, default_keep_segments(overlap) if keep is None else keep)


def write_adj_trace_summary(d: str, out: str, nsteps: int = 200, ckpt: int = 20):
    """Per-schedule mean durations of the adjoint launches and the replay
    stages, from the kernel trace d/run_kernel_trace.csv of
    `bench.py --only adjoint` (launches taken in start order and split by the
    counts each schedule launches: 2 per step for each adjoint kernel, 2 per
    replayed step)."""
    import csv
    import glob
    path = sorted(glob.glob(os.path.join(d, "*kernel_trace.csv")))[0]
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    kinds = {"plasma face VJPs": lambda n: "k_adj_plasma_faces" in n,
             "gather": lambda n: "k_adj_gather" in n,
             "replay stage": lambda n: n.endswith(", 1>(pip::StageArgs)")}
    seqs = {k: [(e - b) / 1e3 for b, e, n in rows if f(n)] for k, f in kinds.items()}
    cmd = (f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --output-format csv -d {d} -o run -- "
           "python3 /root/repo/bench.py --only adjoint")
    lines = ["# Adjoint launches per checkpoint schedule, Alfven-wave damping 512^2, 200 steps: kernel trace",
             "", "Command (from the repository root):", "", f"    {cmd}", "",
             "Mean launch durations (us) of each schedule's gradients, the launches taken in start order and "
             "split by the counts each schedule makes (2 per step for each adjoint kernel, 2 per replayed step). "
             "With the replays beside the adjoint steps the launches of the two streams share the GPU, so "
             "their durations overlap; the in-line schedule gives each launch's standalone duration.", "",
             "| schedule | gradients | plasma face VJPs | gather | replay stage | replay launches |",
             "|---|---|---|---|---|---|"]
    pos = {k: 0 for k in seqs}
    nseg = (nsteps + ckpt - 1) // ckpt
    for _, name, keep, overlap in ADJ_SCHEDULES:
        ng = ADJ_GRADIENTS
        rep = nsteps - _adj_kept(keep, overlap, nseg) * ckpt  # replayed steps (segments of ckpt steps)
        counts = {"plasma face VJPs": 2 * nsteps * ng, "gather": 2 * nsteps * ng, "replay stage": 2 * rep * ng}
        means = {}
        for k, c in counts.items():
            seg = seqs[k][pos[k]:pos[k] + c]
            if len(seg) != c:
                raise SystemExit(f"{path}: {k}: {len(seqs[k])} launches, fewer than the schedules make")
            pos[k] += c
            means[k] = sum(seg) / c
        lines.append(f"| {name} | {ng} | {means['plasma face VJPs']:.1f} | {means['gather']:.1f} | "
                     f"{means['replay stage']:.1f} | {counts['replay stage']} |")
    extra = {k: len(v) - pos[k] for k, v in seqs.items()}
    if any(extra.values()):
        raise SystemExit(f"{path}: launches beyond the schedules: {extra}")
    lines.append("")
    open(out, "w").write("\n".join(lines))
    print("\n".join(lines))


def write_stats_summary(d: str, cmd: str, title: str, out: str, note: str = ""):
    """Kernel-trace summary (rocprofv3 --kernel-trace --stats, csv) as a
    markdown table."""
    rows = _csv_rows(d, "*kernel_stats.csv")
    lines = [f"# {title}", "", "Command (from the repository root):", "", f"    {cmd}", ""]
    if note:
        lines += [note, ""]
    lines += ["| kernel | calls | avg (us) | min (us) | max (us) | share |", "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{r['Name']}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.2f} % |")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def cpu_baseline(wl: str, n: int, steps: int, threads: int):
    """The oracle on `threads` host threads: `steps` consecutive steps of the
    workload, each timed; the median step time gives the rate."""
    from oracle import oracle as orc
    prev = orc.set_threads(threads)
    try:
        cfg, uh, um, ac = make_problem(wl, n)
        o = orc.Oracle(cfg, uh, um, ac)
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            o.step()
            ts.append(time.perf_counter() - t0)
        del o
    finally:
        orc.set_threads(prev)
    med = statistics.median(ts)
    return {"value": cfg.ncells / med, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference Fortran), {WL_NAME[wl]} {n}x{n}, "
                      f"{steps} consecutive steps timed one by one on {threads} host thread(s), median "
                      f"{med:.2f} s per step (steps: {', '.join(f'{t:.2f}' for t in ts)} s)"}


def plain_leg(wl: str, n: int, steps: int, warmup: int, tile_arg: str, arith: str = "exact"):
    """The fused time loop on the workload: wall-clock value and the stage
    kernels' hipEvent timings."""
    import torch
    from pip_amd import autotune
    from pip_amd.solver import PIPSolver
    cfg, uh, um, ac = make_problem(wl, n)
    if tile_arg == "auto":
        tile, _ = autotune.tune_stage(lambda t: PIPSolver(cfg, uh, um, ac, tile=t, arith=arith), cfg, arith=arith)
    else:
        tile = tuple(int(v) for v in tile_arg.split("x"))  # TXxTY[xREMAP[xTWO_PASS_MASK]]
    s = PIPSolver(cfg, uh, um, ac, tile=tile, arith=arith)
    s.enqueue(warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev = s.time_steps(steps)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = s.status()
    s.close()
    assert st == 0, f"stop_sim flags {st}"
    ncell = cfg.ncells
    nv = cfg.nvar_m + cfg.nvar_h
    b_cell = cfg.bytes_per_cell_update()  # 5 nv 8 B (+ 16 B per coefficient field)
    extra = (b_cell - 5 * nv * 8) // 2  # per stage
    b1 = (2 * nv * 8 + extra) * ncell
    b2 = (3 * nv * 8 + extra) * ncell
    t1 = ev["stage1_ms"] / steps * 1e-3
    t2 = ev["stage2_ms"] / steps * 1e-3
    value = ncell * steps / wall
    ach = (b1 + b2) / (t1 + t2) / 1e9
    return cfg, tile, {
        "value": value, "wall_s": wall, "ms_per_step": wall / steps * 1e3,
        "ach_GBs": ach, "b1": b1, "b2": b2, "t1_ms": t1 * 1e3, "t2_ms": t2 * 1e3,
        "b_cell": b_cell, "device_ms_per_step": ev["total_ms"] / steps,
    }


def measured_copy_gbs():
    import torch
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    return 5 * 2 * 8 * (1 << 27) / (e0.elapsed_time(e1) * 1e-3) / 1e9


def batched_leg(n_members: int = 256, tend: float = 0.1, max_steps: int = 20000):
    """Ensemble of n_members 1-D Brio-Wu-type shock tubes (nx log-uniform in
    [256, 4096], xi_n, alpha_c and the pressure ratio drawn from seed 7) run
    to t = tend: the batched ensemble in its graph-captured form (lockstep
    chunks captured as a hipGraph; width autotuned) vs sequential plain
    launches; the autotuner's timings of every mode and width are listed."""
    import torch
    from pip_amd import autotune, problems
    from pip_amd.ensemble import Ensemble
    from pip_amd.solver import PIPSolver
    params = problems.ensemble_members(n_members, seed=7)
    mem = [problems.ensemble_member(p)[:3] for p in params]
    ens = Ensemble(mem, tend=tend, max_steps=max_steps)
    (best_mode, best_tile), tune_ms = autotune.tune_ensemble(ens, f"ens:{n_members}:seed7:tend{tend}")
    mode = "graph"
    tile = min((int(k.split(":")[1]) for k in tune_ms if k.startswith(mode + ":")),
               key=lambda t: tune_ms[f"{mode}:{t}"])
    ens.set_mode(mode)
    ens.set_tile(tile)
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
            "cell_updates": updates, "mode": mode, "tile": tile, "tile_timings_ms": tune_ms,
            "fastest_mode": f"{best_mode}:{best_tile}",
            "ensemble_s": wall_ens, "ensemble_device_ms": ens_ms, "sequential_plain_s": wall_seq,
            "ensemble_value": updates / wall_ens, "sequential_value": updates / wall_seq,
            "unit": "cell-updates/s", "speedup": wall_seq / wall_ens, "bitwise_identical": bool(identical)}


def batched_2d_leg(n_members: int = 64, steps: int = 200):
    """Ensemble of n_members uneven 2-D two-fluid runs (Orszag-Tang states
    with psi, n x n with n log-uniform in [32, 256], xi_n and alpha_c swept;
    seed 9), each for `steps` steps: the graph-captured lockstep ensemble vs
    sequential plain pip_advance runs of the same members."""
    import torch
    from pip_amd import problems
    from pip_amd.ensemble import Ensemble
    from pip_amd.solver import PIPSolver
    mem = [problems.ensemble_member_2d(p)[:3] for p in problems.ensemble_members_2d(n_members)]
    ens = Ensemble(mem, tend=1.0e30, max_steps=steps)
    timing = {}
    for mode in ("lockstep", "graph"):
        ens.set_mode(mode)
        ens.reset()
        ens.run()  # graph mode: captures and caches its graphs
        ens.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.run()
        timing[mode] = time.perf_counter() - t0
    updates, wall_seq, identical = 0, 0.0, True
    for k, (cfg, uh, um) in enumerate(mem):
        s = PIPSolver(cfg, uh, um)
        s.sync()
        t0 = time.perf_counter()
        n, t = s.advance(nsteps=steps)
        wall_seq += time.perf_counter() - t0
        updates += cfg.nx * cfg.ny * n
        r = ens.result(k)
        eh, em = ens.state(k)
        ph, pm = s.state()
        i = (slice(None), slice(2, -2), slice(2, -2))
        identical &= (r["nt"] == n and r["time"] == t and np.array_equal(em[i], pm[i])
                      and np.array_equal(eh[i], ph[i]))
        s.close()
    ens.close()
    sizes = [c.nx for c, _, _ in mem]
    return {"workload": f"{n_members} x 2-D two-fluid Orszag-Tang (psi), n x n with n log-uniform in "
                        f"[32,256] ({min(sizes)}..{max(sizes)}), xi_n and alpha_c swept, {steps} steps each",
            "cell_updates": updates, "mode": "graph", "ensemble_s": timing["graph"],
            "lockstep_s": timing["lockstep"], "sequential_plain_s": wall_seq,
            "ensemble_value": updates / timing["graph"], "sequential_value": updates / wall_seq,
            "unit": "cell-updates/s", "speedup": wall_seq / timing["graph"], "bitwise_identical": bool(identical)}


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
    fwd = []
    for _ in range(3):
        s = PIPSolver(cfg, uh, um, ac, tile=tile)
        s.sync()
        t0 = time.perf_counter()
        s.advance(nsteps)
        fwd.append(time.perf_counter() - t0)
        s.close()
    t_fwd = float(np.median(fwd))
    dev = torch.device("cuda:0")
    U0 = torch.tensor(np.concatenate([um, uh]), device=dev)
    acT = torch.tensor(ac, device=dev)
    mask = torch.zeros((cfg.jx, cfg.ix), dtype=torch.float64, device=dev)
    mask[2:-2, 2:-2] = 1.0

    def gradients(keep, overlap):
        """(recording forward, backward) medians of ADJ_GRADIENTS gradients,
        the first untimed (it allocates the adjoint buffers), and the kept
        segment count"""
        r = TwoFluidRun(cfg, nsteps, ckpt_every=ckpt, tile=tile, keep_segments=keep, overlap_replay=overlap)
        ts = []
        for _ in range(ADJ_GRADIENTS):
            U = U0.clone().requires_grad_(True)
            a = acT.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            UT = r.run(U, a)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            UT.backward(UT.detach() * mask)
            torch.cuda.synchronize()
            ts.append((t1 - t0, time.perf_counter() - t1))
        kept = r.keep_segments
        r.close()
        return float(np.median([t[0] for t in ts[1:]])), float(np.median([t[1] for t in ts[1:]])), kept
    res = {key: gradients(keep, ov) for key, _, keep, ov in ADJ_SCHEDULES}
    t_rec, t_bwd, run_keep = res["default"]
    seg_gb = 4 * ckpt * U0.numel() * 8 / 1e9
    trade = {k: {"gradient_over_forward": (res[k][0] + res[k][1]) / t_fwd, "kept_GB": res[k][2] * seg_gb}
             for k in ("keep_0", "keep_5")}
    in_line = (res["in_line"][0] + res["in_line"][1]) / t_fwd
    return {"workload": f"Alfven-wave damping {n}x{n}, {nsteps} SSP-RK2 steps, checkpoints every {ckpt}",
            "tile": list(tile), "timing": "medians of 3 runs each",
            "forward_s": t_fwd, "forward_recorded_s": t_rec, "backward_s": t_bwd,
            "gradient_over_forward": (t_rec + t_bwd) / t_fwd, "backward_over_forward": t_bwd / t_fwd,
            "kept_segments": run_keep,
            "replay_overlap": {"note": "each replayed segment runs on a second context's stream beside the "
                                       "newer segment's adjoint steps, in two buffer sets that the recording "
                                       "run fills with the 2 newest segments; in_line: the replays in the "
                                       "adjoint's stream, one set, the newest segment kept",
                               "buffer_GB": 2 * seg_gb, "in_line_gradient_over_forward": in_line},
            "memory_for_recompute": {"note": "the same gradient with 0 or 5 of the 10 newest segments' stage "
                                             "inputs kept from the recording run instead of 2 (the rest replayed "
                                             "from their checkpoints; 5 sets hold more memory than the two the "
                                             "sweep needs); not the configuration above",
                                     **trade}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="kh",
                    help="kh: Kelvin-Helmholtz with ionisation/recombination (headline); ot: Orszag-Tang")
    ap.add_argument("--n", type=int, default=0, help="grid size n x n (default: the workload's)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="auto", help="TXxTY[xXCD_ORDER[xTWO_PASS_STAGE_MASK]], or auto (autotuner)")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary, batched and adjoint legs")
    ap.add_argument("--arith", choices=["both", "exact", "relaxed"], default="both",
                    help="stage-kernel arithmetic; both: the exact kernels (bitwise the reference "
                         "path) give the reported value, the relaxed ones the 'other_arith' object")
    ap.add_argument("--pmc-summary", metavar="DIR", help="write the HBM-traffic profile summary from "
                    "rocprofv3 --pmc runs under DIR (fetch, write, copy_fetch, copy_write; the main "
                    "runs used --tile) and exit")
    ap.add_argument("--stats-summary", nargs=3, metavar=("DIR", "OUT", "TITLE"),
                    help="write a kernel-trace summary of the rocprofv3 --stats csv under DIR to OUT and exit")
    ap.add_argument("--sq-summary", nargs=2, metavar=("DIR", "OUT"),
                    help="write the SQ counter summary of the passes under DIR (sq1, sq2) to OUT and exit")
    ap.add_argument("--adj-trace-summary", nargs=2, metavar=("DIR", "OUT"),
                    help="write the per-schedule launch durations of the adjoint leg's kernel trace under DIR "
                         "(rocprofv3 --kernel-trace of --only adjoint) to OUT and exit")
    ap.add_argument("--adj-sq-summary", nargs=2, metavar=("DIR", "OUT"),
                    help="write the adjoint kernels' SQ counter summary of the passes under DIR "
                         "(sq1..sq3 of --only adjoint) to OUT and exit")
    ap.add_argument("--stats-cmd", default="", help="command recorded in --stats-summary")
    ap.add_argument("--stats-note", default="", help="note recorded in --stats-summary")
    ap.add_argument("--only", choices=["batched", "batched_2d", "adjoint", "copy", "ensemble"],
                    help="run one extra leg alone (profiling); copy = 1 GiB device copy that "
                         "calibrates the HBM byte counters; ensemble = the batched leg's ensemble "
                         "alone, twice, in --ens-mode at --ens-tile (no autotuning, no plain runs)")
    ap.add_argument("--ens-mode", choices=["persistent", "lockstep", "graph"], default="graph")
    ap.add_argument("--ens-tile", type=int, choices=[64, 128, 256, 512], default=256)
    args = ap.parse_args()
    n = args.n or WORKLOADS[args.workload][1]
    if args.stats_summary:
        d, out, title = args.stats_summary
        write_stats_summary(d, args.stats_cmd, title, out, args.stats_note)
        return
    if args.adj_sq_summary:
        write_adj_sq_summary(*args.adj_sq_summary)
        return
    if args.adj_trace_summary:
        write_adj_trace_summary(*args.adj_trace_summary)
        return
    if args.sq_summary:
        write_sq_summary(args.sq_summary[0], args.workload, n, args.tile, args.sq_summary[1],
                         "exact" if args.arith == "both" else args.arith)
        return
    if args.pmc_summary:
        pm = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc {c} --output-format csv "
              "-d {d}/{sub} -o run -- python3 /root/repo/bench.py {a}")
        arith = "exact" if args.arith == "both" else args.arith
        main_args = (f"--workload {args.workload} --n {n} --steps 10 --warmup 3 --no-cpu --no-extra "
                     f"--tile {args.tile} --arith {arith}")
        cmds = {sub: pm.format(c=c, d=args.pmc_summary, sub=sub, a=a) for sub, c, a in (
            ("fetch", "FETCH_SIZE", main_args), ("write", "WRITE_SIZE", main_args),
            ("copy_fetch", "FETCH_SIZE", "--only copy"), ("copy_write", "WRITE_SIZE", "--only copy"))}
        write_pmc_summary(args.pmc_summary, cmds, args.workload, n, args.tile, arith)
        return
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU only (asked for {args.gpus}); sharding "
              "across GPUs is designed, not implemented", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: sync around the timed region

    if args.only == "copy":
        nn = 1 << 27  # 2^27 doubles = 1 GiB per buffer
        src = torch.ones(nn, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy": {"bytes_read_per_copy": 8 * nn, "bytes_written_per_copy": 8 * nn}}))
        return
    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_leg(args.ens_mode, args.ens_tile)}))
        return
    if args.only:
        legs = {"batched": batched_leg, "batched_2d": batched_2d_leg, "adjoint": adjoint_leg}
        print(json.dumps({args.only: legs[args.only]()}))
        return

    ariths = ["exact", "relaxed"] if args.arith == "both" else [args.arith]
    legs = {a: plain_leg(args.workload, n, args.steps, args.warmup, args.tile, a) for a in ariths}
    arith = ariths[0]  # the exact kernels: the ones that meet the parity bar
    cfg, tile, r = legs[arith]
    copy_gbs = measured_copy_gbs()
    traffic = traffic_from_profile(args.workload, n, arith)
    config = workload_config(args.workload, n, cfg, tile)
    config.update({"arith": arith, "parity_test": PARITY_TEST[arith]})
    out = {
        "metric": "FP64 two-fluid cell-updates/s",
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic ({WORKLOADS[args.workload][0]} initial state, seeded, no checkpoint)",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": r["ach_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r["ach_GBs"] / HBM_PEAK_GBS,
                     "traffic": traffic["mean"] if traffic else None,
                     "kernel": "k_stage (both SSP-RK stages)",
                     "algorithmic_bytes_per_launch": (r["b1"] + r["b2"]) / 2,
                     "algorithmic_bytes_per_step": r["b1"] + r["b2"],
                     "avg_launch_ms": (r["t1_ms"] + r["t2_ms"]) / 2,
                     "stage1_ms": r["t1_ms"], "stage2_ms": r["t2_ms"],
                     "traffic_stage1": traffic["stage1"] if traffic else None,
                     "traffic_stage2": traffic["stage2"] if traffic else None,
                     "traffic_tile": traffic["tile"] if traffic else None,
                     "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": r["ach_GBs"] / copy_gbs},
        "hbm_roofline_cell_updates_frac": r["value"] * r["b_cell"] / (HBM_PEAK_GBS * 1e9),
        "bytes_per_cell_update": r["b_cell"],
        "device_ms_per_step": r["device_ms_per_step"],
    }
    valu = valu_from_profile(args.workload, n, arith)
    if valu:
        # the bound in practice: FP64 VALU issue (profiles/k_stage_*_sq.md)
        out["valu_ceiling"] = {"lane_instr_per_cell_update": valu[0], "cell_updates_per_s": valu[1],
                               "frac": r["value"] / valu[1]}
    for a in ariths:
        if a == arith:
            continue
        c1, t1, r1 = legs[a]
        tr = traffic_from_profile(args.workload, n, a)
        out["other_arith"] = {"arith": a, "parity_test": PARITY_TEST[a], "tile": list(t1),
                              "value": r1["value"], "ms_per_step": r1["ms_per_step"],
                              "roofline_achieved_GBs": r1["ach_GBs"], "roofline_frac": r1["ach_GBs"] / HBM_PEAK_GBS,
                              "stage1_ms": r1["t1_ms"], "stage2_ms": r1["t2_ms"],
                              "traffic": tr["mean"] if tr else None,
                              "hbm_roofline_cell_updates_frac": r1["value"] * r1["b_cell"] / (HBM_PEAK_GBS * 1e9)}
        v1 = valu_from_profile(args.workload, n, a)
        if v1:
            out["other_arith"]["valu_ceiling"] = {"lane_instr_per_cell_update": v1[0], "cell_updates_per_s": v1[1],
                                                  "frac": r1["value"] / v1[1]}
    if not args.no_extra:
        if args.workload == "kh":
            c2, t2, r2 = plain_leg("ot", 1024, 500, 20, "auto", arith)
            out["secondary"] = {"config": {**workload_config("ot", 1024, c2, t2), "arith": arith}, "steps": 500,
                                "value": r2["value"], "ms_per_step": r2["ms_per_step"],
                                "roofline_achieved_GBs": r2["ach_GBs"],
                                "roofline_frac": r2["ach_GBs"] / HBM_PEAK_GBS,
                                "hbm_roofline_cell_updates_frac": r2["value"] * r2["b_cell"] / (HBM_PEAK_GBS * 1e9)}
        out["batched"] = batched_leg()
        out["batched_2d"] = batched_2d_leg()
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload, n, args.cpu_steps, 1)
        out["cpu_baseline_threads"] = cpu_baseline(args.workload, n, args.cpu_steps,
                                                   min(16, os.cpu_count() or 1))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
