"""GPU checks of the discrete adjoint (pip_amd.adjoint over pip_adj_step).

The gradient of J = 1/2 sum_interior |U_T - U*|^2 with respect to the
per-cell collision coefficient field and the initial velocity is checked
against central finite differences of the frozen-dt replay (the map the
adjoint differentiates), on the 2-D two-fluid Alfven-wave damping problem.

Finite-difference steps are relative (eps = 1e-6): ac_k -> ac_k (1 +- eps)
and v_z,k -> v_z,k +- eps * A with A the wave amplitude (v_z is the wave's
velocity component; v_x = v_y = 0 exactly at the base point, where minmod
has no derivative).  An absolute 1e-6 velocity step would be 1e-3 of the
wave and larger than the second differences the minmod limiter compares
(~1e-7 at 512 cells per wavelength), so the difference quotient would
straddle limiter switches instead of measuring the derivative.  J(+) - J(-) is summed as
sum (U+ - U-) . ((U+ + U-)/2 - U*), which is exact in the cells the
perturbation does not reach.  Agreement is measured on the 64-cell vectors:
the relative L2 error ||g_fd - g_adj|| / ||g_fd|| and the median per-cell
relative error (a cell whose gradient is near zero has a per-cell ratio
dominated by the round-off of the two replays, not by the adjoint).
"""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pip_amd import problems  # noqa: E402
from pip_amd.solver import PIPSolver  # noqa: E402

pytestmark = pytest.mark.gpu

AMP = 1.0e-3


def _setup(n, nsteps, ckpt, seed=11):
    from pip_amd.adjoint import TwoFluidRun
    cfg, uh, um, ac = problems.alfven_damping(n, amp=AMP, seed=seed)
    run = TwoFluidRun(cfg, nsteps, ckpt_every=ckpt)
    dev = torch.device("cuda:0")
    # primitive fields of the initial state (both fluids share v)
    f_n, f_p, f_pn, f_pp = problems.split_fractions(0.5)
    X, Y = problems._grid(cfg)
    base = dict(rho_n=torch.tensor(uh[0], device=dev), rho_p=torch.tensor(um[0], device=dev),
                p_n=torch.full((cfg.jx, cfg.ix), f_pn * 0.1, dtype=torch.float64, device=dev),
                p_p=torch.full((cfg.jx, cfg.ix), f_pp * 0.1, dtype=torch.float64, device=dev),
                B=torch.tensor(um[5:8], device=dev), psi=torch.tensor(um[8], device=dev))
    v0 = torch.tensor(np.stack([um[1] / um[0], um[2] / um[0], um[3] / um[0]]), device=dev)
    return cfg, run, base, v0, torch.tensor(ac, device=dev), (uh, um, ac)


def build_U0(cfg, base, v):
    """pv2cq_mhd / pv2cq_hd (Scheme_rot.f90:30-77) from the shared velocity."""
    g = cfg.gm
    rp, rn = base["rho_p"], base["rho_n"]
    v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    B = base["B"]
    b2 = (B[0] * B[0] + B[1] * B[1]) + B[2] * B[2]
    Um = [rp, rp * v[0], rp * v[1], rp * v[2], (base["p_p"] / (g - 1.0) + (0.5 * rp) * v2) + 0.5 * b2,
          B[0], B[1], B[2], base["psi"]]
    Uh = [rn, rn * v[0], rn * v[1], rn * v[2], base["p_n"] / (g - 1.0) + (0.5 * rn) * v2]
    return torch.stack(Um + Uh).contiguous()


def interior_mask(cfg, dev):
    m = torch.zeros((cfg.jx, cfg.ix), dtype=torch.float64, device=dev)
    m[2:-2, 2:-2] = 1.0
    return m


def target_state(cfg, arrays, nsteps, seed=12):
    """Final snapshot of a run with a different collision field and a 10 %
    larger initial wave (plain fused loop)."""
    uh, um, ac = arrays
    g2 = problems.smooth_field(cfg, seed)
    ac_t = ac * np.exp(0.5 * g2)
    uh_t, um_t = problems.alfven_damping(cfg.nx, amp=1.1 * AMP)[1:3]
    s = PIPSolver(cfg, uh_t, um_t, ac_t)
    s.advance(nsteps)
    uh_T, um_T = s.state()
    s.close()
    return torch.tensor(np.concatenate([um_T, uh_T]), device="cuda:0")


def test_replay_is_bitwise_the_recorded_run():
    print("case replay vs record n=64 steps=45", flush=True)
    cfg, run, base, v0, ac, _ = _setup(64, 45, 7)
    U0 = build_U0(cfg, base, v0)
    out, ckpts, hist = run.forward_record(U0, ac)
    assert len(ckpts) == 7
    rep = run.replay(U0, ac)
    assert torch.equal(out, rep)
    dts = run.dt_history()
    assert np.all(dts > 0) and len(dts) == 45
    run.close()


def test_record_matches_fused_loop():
    """forward_record runs the time loop of pip_advance (same kernels, same
    dt sequence); only the final rewrite round differs (ulp level)."""
    print("case record vs pip_advance n=64 steps=30", flush=True)
    cfg, run, base, v0, ac, (uh, um, acn) = _setup(64, 30, 10)
    out, _, _ = run.forward_record(build_U0(cfg, base, v0), ac)
    s = PIPSolver(cfg, uh, um, acn)
    s.advance(30)
    uh_T, um_T = s.state()
    ref = torch.tensor(np.concatenate([um_T, uh_T]), device="cuda:0")
    m = interior_mask(cfg, ref.device)
    err = ((out - ref).abs() * m).sum() / ((ref.abs() * m).sum())
    assert err < 1e-13, float(err)
    run.close()


def test_saved_primitives_match_recomputed():
    """The replay's saved primitives give bitwise the gradient of the
    adjoint step that recomputes them.  The saving replay (frozen dt, stage
    outputs written into the save buffers) ends bitwise on the recorded
    state and advances time and step count as the recorded steps do."""
    from pip_amd._lib import lib
    from pip_amd.adjoint import _ptr
    print("case saved vs recomputed primitives n=48 steps=12", flush=True)
    grads = []
    for save in (True, False):
        cfg, run, base, v0, ac, _ = _setup(48, 12, 5)
        run.save_primitives = save
        U0 = build_U0(cfg, base, v0)
        out, ckpts, hist = run.forward_record(U0, ac)
        if save:
            bufs = [run._empty_state(12) for _ in range(4)]
            run._start(U0, ac)
            s0 = run.solver.scalars()
            assert lib().pip_replay_steps(run.h, 12, _ptr(run.hist), *[_ptr(b) for b in bufs]) == 0
            s1 = run.solver.scalars()
            rep = run._empty_state()
            assert lib().pip_get_state_device(run.h, _ptr(rep)) == 0
            run.solver.sync()
            assert torch.equal(rep, out)
            t = s0["time"]
            for dt in run.dt_history():
                t = t + float(dt)
            assert s1["time"] == t and s1["nt"] == s0["nt"] + 12
        g = torch.sin(out * 3.0)
        grads.append(run.backward(g, ckpts, ac, hist))
        run.close()
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])


def test_recorded_last_segment_equals_its_replay():
    """The recording run keeps its last segment's stage inputs and their
    primitives (pip_record_steps_saved): they equal what the replay of that
    segment saves, the run ends on the state of a plain recording run, and
    the gradient that skips the last replay equals the one that replays it,
    bit for bit."""
    from pip_amd._lib import lib
    from pip_amd.adjoint import _ptr
    print("case recorded last segment vs replay n=48 steps=12 ckpt=5", flush=True)
    cfg, run, base, v0, ac, _ = _setup(48, 12, 5)
    U0 = build_U0(cfg, base, v0)
    out, ckpts, hist = run.forward_record(U0, ac)
    tag = run.last_run
    assert tag is not None
    kept = [b[:2].clone() for b in run._bufs[0]]
    run.save_primitives = False  # plain pip_record_steps for every segment
    out_plain, _, hist_plain = run.forward_record(U0, ac)
    assert run.last_run is None
    assert torch.equal(out, out_plain) and torch.equal(hist[:24], hist_plain[:24])
    run.save_primitives = True
    bufs = [run._empty_state(2) for _ in range(4)]
    assert lib().pip_set_state_device(run.h, _ptr(ckpts[-1])) == 0
    assert lib().pip_replay_steps(run.h, 2, _ptr(hist) + 16 * 10, *[_ptr(b) for b in bufs]) == 0
    run.solver.sync()
    for a, b in zip(kept, bufs):
        assert torch.equal(a, b)
    g = torch.sin(out * 3.0)
    out2, ckpts2, hist2 = run.forward_record(U0, ac)
    grads_kept = run.backward(g, ckpts2, ac, hist2, run=run.last_run)
    grads_replayed = run.backward(g, ckpts2, ac, hist2)
    assert torch.equal(grads_kept[0], grads_replayed[0]) and torch.equal(grads_kept[1], grads_replayed[1])
    run.close()


def test_kept_segments_give_the_same_gradient():
    """Keeping 0, 1 or 3 of the newest segments from the recording run (the
    rest replayed) gives the same gradient bit for bit; an autograd forward
    on the runner between a forward and its backward makes that backward
    replay every segment."""
    print("case adjoint keep_segments 0 / 1 / 3 n=48 steps=12 ckpt=5 (3 segments)", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    outs = []
    for keep in (0, 1, 3):
        cfg, run0, base, v0, ac, _ = _setup(48, 12, 5)
        run0.close()
        run = TwoFluidRun(cfg, 12, ckpt_every=5, keep_segments=keep)
        U0 = build_U0(cfg, base, v0).requires_grad_(True)
        a = ac.clone().requires_grad_(True)
        UT = run.run(U0, a)
        assert (run.last_run is None) == (keep == 0)
        UT.backward(torch.cos(UT.detach()))
        outs.append((UT.detach(), U0.grad, a.grad))
        if keep == 3:  # a second forward overwrites the kept segments of the first
            U1 = build_U0(cfg, base, v0).requires_grad_(True)
            a1 = ac.clone().requires_grad_(True)
            UTa = run.run(U1, a1)
            UTb = run.run(build_U0(cfg, base, 1.1 * v0), ac)
            UTa.backward(torch.cos(UTa.detach()))
            assert torch.equal(U1.grad, U0.grad) and torch.equal(a1.grad, a.grad)
            del UTb
        run.close()
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert torch.equal(x, y)


def test_overlapped_replays_give_the_in_line_gradient():
    """Replays on the second context's stream beside the adjoint steps
    (overlap_replay, two or more buffer sets reused in turn) give the
    gradient of the in-line sweep bit for bit, with 0, 1 and 3 kept
    segments of 5 (uneven: the last one is shorter), and twice in a row on
    one runner."""
    print("case adjoint overlap_replay vs in line, keep 0 / 1 / 3, n=48 steps=14 ckpt=3 (5 segments)", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    for keep in (0, 1, 3):
        res = []
        for overlap in (False, True):
            cfg, run0, base, v0, ac, _ = _setup(48, 14, 3)
            run0.close()
            run = TwoFluidRun(cfg, 14, ckpt_every=3, keep_segments=keep, overlap_replay=overlap)
            for _ in range(2):
                U0 = build_U0(cfg, base, v0).requires_grad_(True)
                a = ac.clone().requires_grad_(True)
                UT = run.run(U0, a)
                UT.backward(torch.cos(UT.detach()))
                res.append((UT.detach(), U0.grad, a.grad))
            run.close()
        for o in res[1:]:
            for x, y in zip(res[0], o):
                assert torch.equal(x, y), keep


@pytest.mark.parametrize("nsteps,ckpt", [(0, 5), (3, 5), (5, 5), (6, 5)])
def test_short_runs_have_the_in_line_gradient(nsteps, ckpt):
    """Runs of no steps, of one segment shorter than or equal to the
    checkpoint interval, and of a full segment plus one step: the default
    schedule (overlapped replays, two kept segments) gives the in-line
    gradient bit for bit, and with no steps alpha_c has no gradient."""
    print(f"case adjoint short run n=48 steps={nsteps} ckpt={ckpt}", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    res = []
    for overlap in (True, False):
        cfg, run0, base, v0, ac, _ = _setup(48, max(nsteps, 1), ckpt)
        run0.close()
        run = TwoFluidRun(cfg, nsteps, ckpt_every=ckpt, overlap_replay=overlap)
        U0 = build_U0(cfg, base, v0).requires_grad_(True)
        a = ac.clone().requires_grad_(True)
        UT = run.run(U0, a)
        g = torch.cos(UT.detach())
        UT.backward(g)
        res.append((UT.detach(), U0.grad, a.grad))
        if nsteps == 0:
            assert not bool(a.grad.abs().max() > 0)
        run.close()
    for x, y in zip(*res):
        assert torch.equal(x, y)


@pytest.mark.parametrize("bnd", [(1, 1, 1, 1), (2, 2, 1, 1), (1, 1, 3, 2)])
def test_adjoint_step_equals_host_work_items(bnd):
    """One adjoint step through the C-ABI (pip_adj_seed, pip_adj_step) equals
    the host build of the work items run one after another (hc_adj_step, the
    step the CPU tests check against central differences of the oracle), on
    a 37 x 29 grid that leaves partial gather tiles on both edges; periodic
    and mirror boundaries.  The kernels keep the neutrals' face adjoints of a
    tile in shared memory and the host writes them to planes; the device
    VJPs divide through a refined reciprocal and contract the multiply-adds
    of their reverse sweep (the host divides exactly and rounds every
    operation), so the bar is 1e-13 of the largest entry.  It holds at the
    mirror walls, where S_M = 0 exactly by symmetry, because the VJPs'
    restated forward is not contracted, and at small contact speeds, where
    the derivative of HLLC's smoothed switch is zero, because it is not
    formed as a difference of quotients."""
    from oracle import oracle as orc
    from pip_amd._lib import check, lib
    from pip_amd.config import PIPConfig
    from tests.test_adjoint_cpu import L as HOST, _smooth_state
    print(f"case adjoint step vs host work items 37x29 flag_bnd={bnd}", flush=True)
    cfg = PIPConfig(ndim=2, nx=37, ny=29, xmin=0.0, xmax=1.0, ymin=0.0, ymax=29 / 37, flag_divb=1,
                    ac_field=True, flag_bnd=bnd, col=30.0)
    uh, um, ac = _smooth_state(cfg, 5)
    o = orc.Oracle(cfg, uh, um, ac, prologue=False)
    o.pipbnd()
    uh, um = o.state()
    dt = 1.0e-3
    cmax = cfg.safety * min(cfg.dx, cfg.dy) / dt
    o.set_scalars(dt, cmax)
    U1 = np.ascontiguousarray(o.run_solver_stage1())
    Un = np.ascontiguousarray(np.concatenate([um, uh]))
    rng = np.random.default_rng(17)
    W = rng.normal(size=Un.shape)  # ghost entries too: the seed folds them
    lam_h, lam_ac_h = W.copy(), np.zeros((cfg.jx, cfg.ix))
    P = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    HOST.hc_adj_step.argtypes = [I, I, P, ctypes.POINTER(I), P, P, P, P, P]
    prm = np.array([cfg.gm, cfg.ro_lim, cfg.pr_lim, cfg.dx, cfg.dy, cfg.col, cfg.safety, cfg.db_clean, dt, cmax])
    HOST.hc_adj_step(cfg.ix, cfg.jx, prm.ctypes.data_as(P), (I * 4)(*bnd), Un.ctypes.data_as(P),
                     U1.ctypes.data_as(P), np.ascontiguousarray(ac).ctypes.data_as(P), lam_h.ctypes.data_as(P),
                     lam_ac_h.ctypes.data_as(P))
    s = PIPSolver(cfg, uh, um, ac, prologue=False)
    dev = torch.device("cuda:0")
    Un_d, U1_d = torch.tensor(Un, device=dev), torch.tensor(U1, device=dev)
    dtc = torch.tensor([dt, cmax], dtype=torch.float64, device=dev)
    lam = torch.tensor(W, device=dev)
    lam_ac = torch.zeros((cfg.jx, cfg.ix), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    check(lib().pip_adj_seed(s.h, lam.data_ptr()))
    check(lib().pip_adj_step(s.h, Un_d.data_ptr(), U1_d.data_ptr(), None, None, dtc.data_ptr(), lam.data_ptr(),
                             lam_ac.data_ptr()))
    s.sync()
    g, g_ac = lam.cpu().numpy(), lam_ac.cpu().numpy()
    s.close()
    scale, scale_ac = np.abs(lam_h).max(), np.abs(lam_ac_h).max()
    err, err_ac = np.abs(g - lam_h).max() / scale, np.abs(g_ac - lam_ac_h).max() / scale_ac
    print(f"  max |gpu - host| / max |host|: state {err:.2e}, alpha {err_ac:.2e}", flush=True)
    assert scale > 0 and scale_ac > 0
    assert err <= 1e-13 and err_ac <= 1e-13
    assert np.all(g[:, :2] == 0) and np.all(g[:, -2:] == 0) and np.all(g[:, :, :2] == 0) and np.all(g[:, :, -2:] == 0)


def test_two_pass_stage_kernels_give_the_same_gradient():
    """Forward, replay (with primitive saves) and gradient through the
    fluid-by-fluid stage kernels equal the one-pass kernels' bit for bit."""
    print("case adjoint two-pass vs one-pass n=48 steps=12", flush=True)
    outs = []
    for tile in ((16, 15, 1, 0), (16, 15, 1, 3)):
        cfg, run0, base, v0, ac, _ = _setup(48, 12, 5)
        run0.close()
        from pip_amd.adjoint import TwoFluidRun
        run = TwoFluidRun(cfg, 12, ckpt_every=5, tile=tile, replay_tile=tile)
        U0 = build_U0(cfg, base, v0)
        out, ckpts, hist = run.forward_record(U0, ac)
        lam, lam_ac = run.backward(torch.cos(out), ckpts, ac, hist)
        outs.append((out, lam, lam_ac))
        run.close()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_replay_tile_differs_from_the_recording_tile():
    """Recording with the fluid-by-fluid kernels and replaying one-pass (the
    default for a two-pass tile) gives the gradient of an all-one-pass run."""
    print("case adjoint record 16x15x1x3, replay auto n=48 steps=12", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    outs = []
    for tile, rt in (((16, 15, 1, 0), (16, 15, 1, 0)), ((16, 15, 1, 3), "auto")):
        cfg, run0, base, v0, ac, _ = _setup(48, 12, 5)
        run0.close()
        run = TwoFluidRun(cfg, 12, ckpt_every=5, tile=tile, replay_tile=rt)
        assert run.replay_tile == (16, 15, 1, 0)
        out, ckpts, hist = run.forward_record(build_U0(cfg, base, v0), ac)
        outs.append((out,) + run.backward(torch.cos(out), ckpts, ac, hist))
        run.close()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_two_forwards_then_one_backward():
    """Each differentiable run keeps its own recorded dt sequence: two
    forwards on one runner (different alpha fields, hence different dt
    sequences), then the first one's backward, give the gradient of the
    first run done alone."""
    print("case adjoint two forwards, one backward n=48 steps=12", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    cfg, run, base, v0, ac, _ = _setup(48, 12, 5)
    U0 = build_U0(cfg, base, v0)

    def grads(runner, a, others=()):
        U = U0.clone().requires_grad_(True)
        av = a.clone().requires_grad_(True)
        UT = runner.run(U, av)
        for o in others:  # later forwards on the same runner
            runner.run(U0.clone().requires_grad_(True), o.clone().requires_grad_(True))
        (UT * torch.cos(UT)).sum().backward()
        return UT.detach(), U.grad, av.grad

    ref = grads(run, ac)
    run.close()
    run = TwoFluidRun(cfg, 12, ckpt_every=5)
    got = grads(run, ac, others=(2.0 * ac,))
    h_other = run.dt_history()
    run.close()
    run = TwoFluidRun(cfg, 12, ckpt_every=5)
    run.run(U0.clone(), ac.clone())
    assert not np.array_equal(run.dt_history(), h_other)  # the second run's dt sequence differs
    run.close()
    for a, b in zip(ref, got):
        assert torch.equal(a, b)


def _fd_check(n, nsteps, ckpt, ncells, tol, seed=13, eps=1.0e-6, eps_ac=None):
    cfg, run, base, v0, ac, arrays = _setup(n, nsteps, ckpt)
    dev = ac.device
    Ustar = target_state(cfg, arrays, nsteps)
    mask = interior_mask(cfg, dev)
    v = v0.clone().requires_grad_(True)
    acv = ac.clone().requires_grad_(True)
    UT = run.run(build_U0(cfg, base, v), acv)
    J = 0.5 * (((UT - Ustar) ** 2) * mask).sum()
    J.backward()
    g_ac, g_v = acv.grad.clone(), v.grad.clone()
    # gradient support: interior only
    assert float(g_ac[:2].abs().max()) == 0.0 and float(g_ac[:, :2].abs().max()) == 0.0

    def dJ(Up, Um):
        return float((((Up - Um) * (0.5 * (Up + Um) - Ustar)) * mask).sum())

    rng = np.random.default_rng(seed)
    cells = [(int(rng.integers(2, cfg.jx - 2)), int(rng.integers(2, cfg.ix - 2))) for _ in range(ncells)]
    eps_ac = eps if eps_ac is None else eps_ac
    fd_ac, ad_ac, fd_v, ad_v = [], [], [], []
    with torch.no_grad():
        U0 = build_U0(cfg, base, v0)
        for c, (j, i) in enumerate(cells):
            if c % 16 == 0:
                print(f"  FD cell {c}/{len(cells)}", flush=True)
            h = eps_ac * float(ac[j, i])
            acp, acm = ac.clone(), ac.clone()
            acp[j, i] += h
            acm[j, i] -= h
            fd_ac.append(dJ(run.replay(U0, acp), run.replay(U0, acm)) / (2 * h))
            ad_ac.append(float(g_ac[j, i]))
            # the wave component; v_x = v_y = 0 exactly at the base point,
            # which puts their limiters at minmod's kink (no derivative)
            h = eps * AMP
            vp, vm = v0.clone(), v0.clone()
            vp[2, j, i] += h
            vm[2, j, i] -= h
            fd_v.append(dJ(run.replay(build_U0(cfg, base, vp), ac), run.replay(build_U0(cfg, base, vm), ac)) / (2 * h))
            ad_v.append(float(g_v[2, j, i]))
    run.close()
    res = {}
    for name, fd, ad in (("ac", fd_ac, ad_ac), ("v_z", fd_v, ad_v)):
        fd, ad = np.array(fd), np.array(ad)
        agg = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
        per = np.abs(fd - ad) / np.abs(fd)
        res[name] = (agg, np.median(per), per.max())
        print(f"FD check n={n} steps={nsteps} eps={eps if name == 'v_z' else eps_ac:g} {name}: aggregate "
              f"rel L2 {agg:.3e}, per-cell median {np.median(per):.3e}, max {per.max():.3e}", flush=True)
    tol_med, tol_agg = tol
    for name in ("ac", "v_z"):
        assert res[name][1] <= tol_med and res[name][0] <= tol_agg, (name, res[name])
    return res


def test_adjoint_fd_small():
    """48^2, 36 steps, checkpoints every 7 (uneven last segment), 16 cells:
    per-cell median relative error <= 1e-6; the vector's relative L2 error
    <= 1e-5 (one cell whose perturbation region holds a near-tie of the
    limiter can carry ~1e-5 on its own)."""
    print("case adjoint FD n=48 steps=36 ckpt=7", flush=True)
    _fd_check(48, 36, 7, 16, (1e-6, 1e-5))


def _gradient_512():
    cfg, run, base, v0, ac, arrays = _setup(512, 200, 20)
    dev = ac.device
    Ustar = target_state(cfg, arrays, 200)
    mask = interior_mask(cfg, dev)
    v = v0.clone().requires_grad_(True)
    acv = ac.clone().requires_grad_(True)
    UT = run.run(build_U0(cfg, base, v), acv)
    J = 0.5 * (((UT - Ustar) ** 2) * mask).sum()
    J.backward()
    return cfg, run, base, v0, ac, Ustar, mask, acv.grad.clone(), v.grad.clone()


def _directional(cfg, run, base, v0, ac, Ustar, mask, g_ac, g_v, d, eps):
    """(FD, adjoint) of <dJ/d(log ac), d> and of <dJ/dv_z, A d>."""
    def dJ(Up, Um):
        return float((((Up - Um) * (0.5 * (Up + Um) - Ustar)) * mask).sum())
    with torch.no_grad():
        U0 = build_U0(cfg, base, v0)
        fd_a = dJ(run.replay(U0, ac * (1 + eps * d)), run.replay(U0, ac * (1 - eps * d))) / (2 * eps)
        vp, vm = v0.clone(), v0.clone()
        vp[2] += eps * AMP * d
        vm[2] -= eps * AMP * d
        fd_v = dJ(run.replay(build_U0(cfg, base, vp), ac), run.replay(build_U0(cfg, base, vm), ac)) / (2 * eps)
    return fd_a, float((g_ac * ac * d).sum()), fd_v, float((g_v[2] * AMP * d).sum())


# Width (cells) of the smooth perturbation directions at 512^2.  The map is
# piecewise smooth: in a resolved wave the two slopes minmod compares differ
# by ~1e-7 of the wave, so a perturbation whose own second differences are
# comparable flips limiter branches and the difference quotient picks up the
# jumps of the derivative there.  Directions smooth on the grid scale barely
# move those margins.  An eps sweep (1e-8 .. 1e-3, four seeds) with
# Gaussian-filtered noise of width ~25 cells gives FD/adjoint agreement
# within 1e-7 for eps = 1e-7 .. 1e-6 and within 4e-7 up to 1e-4 for both
# parameters; with width ~1.3 cells the error grows in proportion to eps
# (branch flips, up to 5e-3 at eps = 1e-4) and, below eps ~ 1e-7, as 1/eps
# (round-off).
SMOOTH_FILT = 160.0   # problems.smooth_field filter: width 160 / (2 pi) ~ 25 cells
BUMP_SIGMA = 25.0     # cells


def test_adjoint_directional_fd_alfven_512():
    """512x512 periodic, 200 SSP-RK2 steps, checkpoints every 20: the
    adjoint's directional derivatives <g, d> against central differences of
    the replay along four smooth random directions d over all interior cells
    (unit-variance Gaussian-filtered noise of width ~25 cells, seeds 21-24):
    the alpha_c field as ac (1 +- eps d) and the initial v_z as v_z +- eps A d
    (A the wave amplitude), eps = 1e-6.  Relative error <= 1e-6 for each."""
    print("case adjoint directional FD n=512 steps=200 ckpt=20", flush=True)
    cfg, run, base, v0, ac, Ustar, mask, g_ac, g_v = _gradient_512()
    dev = ac.device
    worst = 0.0
    for seed in (21, 22, 23, 24):
        d = torch.tensor(problems.smooth_field(cfg, seed, filt=SMOOTH_FILT), device=dev) * mask
        fd_a, ad_a, fd_v, ad_v = _directional(cfg, run, base, v0, ac, Ustar, mask, g_ac, g_v, d, 1.0e-6)
        ea, ev = abs(fd_a - ad_a) / abs(fd_a), abs(fd_v - ad_v) / abs(fd_v)
        print(f"directional FD seed {seed} eps=1e-6: ac {fd_a:.10e} vs {ad_a:.10e} (rel {ea:.2e}); "
              f"v_z {fd_v:.10e} vs {ad_v:.10e} (rel {ev:.2e})", flush=True)
        worst = max(worst, ea, ev)
    run.close()
    assert worst <= 1e-6, worst


def test_adjoint_local_bumps_fd_alfven_512():
    """The same run, 64 random cells (seed 13): at each, a Gaussian bump of
    width 25 cells centred on the cell (periodic distance) as the
    perturbation direction of alpha_c (relative) and of v_z (times the wave
    amplitude), eps = 1e-6.  This samples the gradient around 64 places of
    the box with perturbations smooth enough for central differences to
    resolve the derivative (one-cell perturbations are not: see
    test_adjoint_fd_alfven_512).  Relative error per bump <= 1e-6, median
    <= 1e-7.  (With 12-cell bumps a few alpha_c bumps straddle a sign change
    of the gradient, |sum g d| ~ 1e-2 sum |g d|, and their relative error
    reaches ~1e-5 while the error relative to sum |g d| stays <= 3e-7.)"""
    print("case adjoint local-bump FD n=512 steps=200 ckpt=20, 64 cells", flush=True)
    cfg, run, base, v0, ac, Ustar, mask, g_ac, g_v = _gradient_512()
    dev = ac.device
    rng = np.random.default_rng(13)
    cells = [(int(rng.integers(2, cfg.jx - 2)), int(rng.integers(2, cfg.ix - 2))) for _ in range(64)]
    jj = torch.arange(cfg.jx, device=dev, dtype=torch.float64)[:, None]
    ii = torch.arange(cfg.ix, device=dev, dtype=torch.float64)[None, :]
    ea_all, ev_all = [], []
    for c, (j, i) in enumerate(cells):
        dj = torch.remainder(jj - j + cfg.ny / 2, cfg.ny) - cfg.ny / 2
        di = torch.remainder(ii - i + cfg.nx / 2, cfg.nx) - cfg.nx / 2
        d = torch.exp(-(dj * dj + di * di) / (2 * BUMP_SIGMA ** 2)) * mask
        fd_a, ad_a, fd_v, ad_v = _directional(cfg, run, base, v0, ac, Ustar, mask, g_ac, g_v, d, 1.0e-6)
        ea_all.append(abs(fd_a - ad_a) / abs(fd_a))
        ev_all.append(abs(fd_v - ad_v) / abs(fd_v))
        if c % 16 == 15:
            print(f"  bumps {c + 1}/64: max so far ac {max(ea_all):.2e}, v_z {max(ev_all):.2e}", flush=True)
    run.close()
    ea_all, ev_all = np.array(ea_all), np.array(ev_all)
    print(f"local-bump FD n=512 eps=1e-6: ac median {np.median(ea_all):.2e} max {ea_all.max():.2e}; "
          f"v_z median {np.median(ev_all):.2e} max {ev_all.max():.2e}", flush=True)
    assert ea_all.max() <= 1e-6 and ev_all.max() <= 1e-6
    assert np.median(ea_all) <= 1e-7 and np.median(ev_all) <= 1e-7


def test_adjoint_fd_alfven_512():
    """512x512 periodic, 200 SSP-RK2 steps, checkpoints every 20 steps,
    64 random cells (seed 13), one cell perturbed at a time.

    At this size a one-cell central difference resolves the derivative only
    to ~1e-6-1e-5.  A sweep of eps from 1e-7 to 3e-3 on twelve cells shows
    no plateau below ~1e-6: the signal of one cell's perturbation scales with
    the cell area while the replays' rounding differences grow with the
    region it reaches (the error falls as 1/eps up to ~1e-5), and larger
    steps cross limiter near-ties -- with 360 cells per wavelength the
    slopes minmod compares differ by ~1e-7 of the wave -- where the map has
    no derivative (the error grows with eps beyond ~1e-4).  The alpha_c step
    is therefore 3e-5 (relative; the middle of the sweep's flat part), the
    v_z step 1e-6 of the amplitude, and the bars are what such differences
    resolve: per-cell median <= 2e-5, relative L2 <= 1e-4.  The derivative
    itself is held to 1e-6 at this size by test_adjoint_directional_fd_alfven_512
    and test_adjoint_local_bumps_fd_alfven_512 (smooth perturbations), to
    1e-6 per cell at 48^2, and to 1e-8 by the one-step check against the
    oracle's differences."""
    print("case adjoint FD n=512 steps=200 ckpt=20", flush=True)
    _fd_check(512, 200, 20, 64, (2e-5, 1e-4), eps_ac=3e-5)
