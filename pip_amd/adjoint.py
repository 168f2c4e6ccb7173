"""Differentiable two-fluid runs: a torch.autograd.Function over the adjoint
entry points of libpipgpu (pip_record_steps, pip_replay_steps, pip_adj_seed,
pip_adj_step).

``TwoFluidRun(cfg, nsteps, ckpt_every=20)`` integrates ``nsteps`` steps of
the time loop (main.f90:31-68: prologue, then fused set_IR/cfl/run_solver
steps) from a device state ``U0`` (nvar_m + 5, jx, ix) with a per-cell
collision coefficient field ``ac`` (jx, ix).  ``run(U0, ac)`` returns the
final state and is differentiable with respect to both inputs:

  forward   records dt/cmax of every step on the device and keeps the state
            at every ``ckpt_every``-th step (step checkpointing); the newest
            segments, as many as the backward sweep has buffer sets (two by
            default), also keep their stage inputs (what their replays
            would save);
  backward  for each segment, newest first: restore its checkpoint, replay
            the segment with the recorded dt (bitwise the forward run)
            saving the stage inputs (the newest segments' are still there
            from the forward), then pull the adjoint back through its steps
            (hand-written reverse-mode kernels: stage update, HLLD/HLLC
            fluxes, minmod reconstruction, collision source).  By default
            the replay of the next older segment runs on a second library
            context's stream while the adjoint steps of the current one run,
            into the other of two buffer sets (replay_sets).

The gradient is that of the discrete map with the recorded time-step
sequence held fixed (dt and the GLM speed cmax are treated as constants);
``replay(U0, ac)`` evaluates exactly that map, so central differences of
``replay`` check the gradient.  The library is the only compute path: with
no built library every call raises.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .config import PIPConfig
from ._lib import check, lib
from .solver import PIPSolver


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def default_keep_segments(overlap_replay: bool) -> int:
    """Segments the recording run keeps by default: as many as the backward
    sweep has buffer sets (two with overlapped replays, one in line)."""
    return 2 if overlap_replay else 1


def replay_sets(nseg: int, kept: int, overlap: bool):
    """Buffer sets of the backward sweep over nseg segments, i-th newest
    first: (sets, waits).  Segment i reads set sets[i]; the kept newest
    segments (i < kept) hold sets 0..kept-1 from the recording run.  In line
    (overlap False) every replayed segment uses set 0.  Overlapped, the
    replay of segment i + 1 runs while the adjoint steps of segment i read
    their set, so neighbours take different sets: a replayed segment reuses
    the set of segment i - 2, and waits[i] names the last segment that read
    it before (the replay waits for its adjoint steps; -1: none)."""
    if nseg < 0 or kept < 0 or kept > nseg:
        raise ValueError("0 <= kept <= nseg")
    sets = []
    for i in range(nseg):
        if i < kept:
            sets.append(i)
        elif not overlap:
            sets.append(0)
        else:
            sets.append(i if i < 2 else sets[i - 2])
    waits = []
    for i in range(nseg):
        prev = [j for j in range(i) if sets[j] == sets[i]]
        waits.append(prev[-1] if (overlap and i >= kept and prev) else -1)
    return sets, waits


class TwoFluidRun:
    def __init__(self, cfg: PIPConfig, nsteps: int, ckpt_every: int = 20, tile=None,
                 device: str = "cuda:0", save_primitives: bool = True, replay_tile="auto",
                 keep_segments: int | None = None, overlap_replay: bool = True):
        """keep_segments: how many of the newest segments the recording run
        keeps the stage inputs of (4 states per step each, 2.4 GB per
        20-step segment at 512^2), so that the backward sweep replays only
        the others; 0 replays every segment.  overlap_replay: the backward
        sweep replays each segment on a second context's stream while the
        adjoint steps of the newer one run (bitwise the same gradient; at
        least two buffer sets); False replays in line, into one set.  The
        default keep_segments fills the buffer sets the backward sweep holds
        anyway: 2 with overlap_replay, 1 without."""
        if cfg.ndim != 2 or cfg.nvar_m != 9 or cfg.flag_ir != 0 or cfg.t_order != 2:
            raise ValueError("the adjoint covers 2-D, flag_divb=1, flag_IR=0, SSP-RK2")
        if not cfg.ac_field:
            raise ValueError("the adjoint run needs a per-cell ac field (cfg.ac_field)")
        if keep_segments is None:
            keep_segments = default_keep_segments(overlap_replay)
        if nsteps < 0 or ckpt_every < 1 or keep_segments < 0:
            raise ValueError("nsteps >= 0, ckpt_every >= 1 and keep_segments >= 0")
        self.cfg = cfg
        self.nsteps = int(nsteps)
        self.ckpt_every = int(ckpt_every)
        self.device = torch.device(device)
        self.nvt = cfg.nvar_m + 5
        z = np.zeros((cfg.jx, cfg.ix))
        self.solver = PIPSolver(cfg, np.zeros((5, cfg.jx, cfg.ix)), np.zeros((cfg.nvar_m, cfg.jx, cfg.ix)),
                                z, prologue=False, tile=tile)
        # (tx, ty, xcd order, two-pass stage mask) of the recording run, and of
        # the replays: every tile gives the same bits, and the replay kernels
        # (no dt reduction) run fastest one-pass on 16x15 tiles in the
        # recording run's workgroup order (the default "auto"; measured at
        # 512^2 in line and beside the adjoint steps)
        t = list(tile) if tile is not None else []
        self.tile = tuple(t[:2] + [t[2] if len(t) > 2 else None, t[3] if len(t) > 3 else 0]) if t else None
        if isinstance(replay_tile, str) and replay_tile == "auto":
            replay_tile = (16, 15, self.tile[2], 0) if self.tile else None
        self.replay_tile = tuple(replay_tile) if replay_tile is not None else None
        self.h = self.solver.h
        self.hist = torch.zeros(2 * max(self.nsteps, 1), dtype=torch.float64, device=self.device)
        # segment buffers: stage inputs and their primitives, 4 states per
        # step; buffer set i holds the i-th newest kept segment, and set 0
        # also takes the replays of the others
        self.keep_segments = int(keep_segments)
        self._bufs = []
        # the forward run whose kept segments the buffers hold (None: none)
        self._seg_run = None
        self.last_run = None
        # the replay also saves the stage inputs' primitives (twice the segment
        # memory, no primitive pass in the adjoint step)
        self.save_primitives = bool(save_primitives)
        self.overlap_replay = bool(overlap_replay)
        self._rsolver = None  # the replay context (overlap_replay), made on first use

    # ------------------------------------------------------------ helpers
    def _segments(self):
        out = []
        k = 0
        while k < self.nsteps:
            m = min(self.ckpt_every, self.nsteps - k)
            out.append((k, m))
            k += m
        return out

    def _torch_done(self):
        torch.cuda.current_stream(self.device).synchronize()

    def _lib_done(self):
        self.solver.sync()

    def _use_tile(self, tile):
        if tile is not None:
            self.solver.set_tile(*tile)

    def _start(self, U0: torch.Tensor, ac: torch.Tensor):
        """Load (U0, ac), run the prologue (Initial_rot.f90:140-160) and the
        first loop-top set_IR/cfl round (main.f90:32-36)."""
        self._check_inputs(U0, ac)
        self._torch_done()
        L = lib()
        check(L.pip_set_state_device(self.h, _ptr(U0)))
        check(L.pip_set_ac_device(self.h, _ptr(ac)))
        dt = ctypes.c_double()
        check(L.pip_prologue(self.h, ctypes.byref(dt)))
        check(L.pip_record_steps(self.h, 0, None))

    def _check_inputs(self, U0, ac):
        c = self.cfg
        if tuple(U0.shape) != (self.nvt, c.jx, c.ix) or U0.dtype != torch.float64 or not U0.is_contiguous():
            raise ValueError(f"U0 must be a contiguous float64 tensor of shape {(self.nvt, c.jx, c.ix)}")
        if tuple(ac.shape) != (c.jx, c.ix) or ac.dtype != torch.float64 or not ac.is_contiguous():
            raise ValueError(f"ac must be a contiguous float64 tensor of shape {(c.jx, c.ix)}")
        if U0.device != self.device or ac.device != self.device:
            raise ValueError(f"inputs must live on {self.device}")

    def _kept(self):
        """Number of newest segments the recording run keeps."""
        return min(self.keep_segments, len(self._segments())) if self.save_primitives else 0

    def _segment_buffers(self, nsets: int):
        """The first nsets buffer sets (Un, U1, Vn, V1: m states each; Vn, V1
        None without primitive saves), allocated on first use."""
        mmax = max((m for _, m in self._segments()), default=0)
        if mmax and self._bufs and self._bufs[0][0].shape[0] < mmax:
            self._bufs = []
        while mmax and len(self._bufs) < nsets:
            sv = self.save_primitives
            self._bufs.append((self._empty_state(mmax), self._empty_state(mmax),
                               self._empty_state(mmax) if sv else None, self._empty_state(mmax) if sv else None))
        return self._bufs

    def _empty_state(self, n=None):
        c = self.cfg
        shape = (self.nvt, c.jx, c.ix) if n is None else (n, self.nvt, c.jx, c.ix)
        return torch.empty(shape, dtype=torch.float64, device=self.device)

    # ------------------------------------------------------------ passes
    def forward_record(self, U0: torch.Tensor, ac: torch.Tensor):
        """Forward run recording dt/cmax; returns (U_T, checkpoints, history),
        history being this run's own copy of the recorded (dt, cmax) pairs
        (the runner's buffer is overwritten by the next forward).  The kept
        newest segments run through pip_record_steps_saved, which leaves
        their stage inputs in the segment buffers; ``last_run`` names this
        run for backward() until another forward or a replay overwrites
        them."""
        self._start(U0, ac)
        L = lib()
        ckpts = []
        segs = self._segments()
        kept = self._kept()
        bufs = self._segment_buffers(kept)
        self._seg_run = None
        for s_, (k0, m) in enumerate(segs):
            ck = self._empty_state()
            check(L.pip_get_state_device(self.h, _ptr(ck)))
            ckpts.append(ck)
            i = len(segs) - 1 - s_  # i-th newest segment
            if i < kept:
                check(L.pip_record_steps_saved(self.h, m, _ptr(self.hist) + 16 * k0,
                                               *[_ptr(b) for b in bufs[i]]))
            else:
                check(L.pip_record_steps(self.h, m, _ptr(self.hist) + 16 * k0))
        if kept:
            self._seg_run = object()
        self.last_run = self._seg_run
        out = self._empty_state()
        check(L.pip_get_state_device(self.h, _ptr(out)))
        self._lib_done()
        return out, ckpts, self.hist.clone()

    def replay(self, U0: torch.Tensor, ac: torch.Tensor, hist: torch.Tensor | None = None) -> torch.Tensor:
        """The same run with a recorded dt/cmax sequence held fixed (the map
        whose derivative backward() returns); hist defaults to the last
        recorded run's."""
        hist = self.hist if hist is None else hist
        self._start(U0, ac)
        self._torch_done()
        self._use_tile(self.replay_tile)
        check(lib().pip_replay_steps(self.h, self.nsteps, _ptr(hist), None, None, None, None))
        self._use_tile(self.tile)
        out = self._empty_state()
        check(lib().pip_get_state_device(self.h, _ptr(out)))
        self._lib_done()
        return out

    def backward(self, gU: torch.Tensor, ckpts, ac: torch.Tensor, hist: torch.Tensor, run=None):
        """dJ/dU0, dJ/dac from dJ/dU_T by checkpointed reverse sweeps over the
        forward run that produced ckpts and hist (``run``: its ``last_run``;
        when the segment buffers still hold that run's kept segments, their
        replays are skipped)."""
        if hist.numel() < 2 * self.nsteps or len(ckpts) != len(self._segments()):
            raise ValueError("checkpoints and history do not belong to a run of this runner")
        L = lib()
        lam = gU.detach().to(torch.float64).contiguous().clone()
        lam_ac = torch.zeros_like(ac)
        segs = self._segments()
        kept = self._kept() if (run is not None and run is self._seg_run) else 0
        overlap = self.overlap_replay and len(segs) > kept
        sets, waits = replay_sets(len(segs), kept, overlap)
        bufs = self._segment_buffers(max(sets, default=-1) + 1)
        self._torch_done()
        check(L.pip_set_ac_device(self.h, _ptr(ac)))
        check(L.pip_adj_seed(self.h, _ptr(lam)))
        sb = self.nvt * self.cfg.ix * self.cfg.jx * 8  # bytes per saved state
        rsegs = list(zip(reversed(segs), reversed(ckpts)))  # i-th newest first
        if overlap:
            rs = self._replay_solver()
            check(L.pip_set_ac_device(rs.h, _ptr(ac)))
            rctx = rs.h
            main_s = torch.cuda.ExternalStream(self.solver.stream(), device=self.device)
            rep_s = torch.cuda.ExternalStream(rs.stream(), device=self.device)
        else:
            rctx = self.h
            self._use_tile(self.replay_tile)
        adj_done, rep_done = {}, {}

        def replay(i):
            # segment i from its checkpoint into buffer set sets[i]; with the
            # overlap, after the adjoint steps of the last segment that read
            # that set (waits[i]) and on the replay context's stream
            (k0, m), ck = rsegs[i]
            Un, U1, Vn, V1 = bufs[sets[i]]
            self._seg_run = None  # a replay overwrites the kept buffers
            if overlap and waits[i] >= 0:
                rep_s.wait_event(adj_done[waits[i]])
            check(L.pip_set_state_device(rctx, _ptr(ck)))
            check(L.pip_replay_steps(rctx, m, _ptr(hist) + 16 * k0, _ptr(Un), _ptr(U1),
                                     _ptr(Vn) if Vn is not None else None,
                                     _ptr(V1) if V1 is not None else None))
            if overlap:
                rep_done[i] = torch.cuda.Event()
                rep_done[i].record(rep_s)

        try:
            if overlap and kept == 0:
                replay(0)
            for i, ((k0, m), _) in enumerate(rsegs):
                if overlap:
                    if i + 1 < len(rsegs) and i + 1 >= kept:
                        replay(i + 1)  # runs beside this segment's adjoint steps
                    if i >= kept:
                        main_s.wait_event(rep_done[i])
                elif i >= kept:
                    replay(i)
                Un, U1, Vn, V1 = bufs[sets[i]]
                for k in reversed(range(m)):
                    check(L.pip_adj_step(self.h, _ptr(Un) + sb * k, _ptr(U1) + sb * k,
                                         _ptr(Vn) + sb * k if Vn is not None else None,
                                         _ptr(V1) + sb * k if V1 is not None else None,
                                         _ptr(hist) + 16 * (k0 + k), _ptr(lam), _ptr(lam_ac)))
                if overlap:
                    adj_done[i] = torch.cuda.Event()
                    adj_done[i].record(main_s)
        finally:
            if overlap:
                if self._rsolver is not None:
                    self._rsolver.sync()
            else:
                self._use_tile(self.tile)
            self._lib_done()
        return lam, lam_ac

    def _replay_solver(self) -> PIPSolver:
        """The second library context the overlapped replays run on (its own
        stream and state; the replay tile)."""
        if self._rsolver is None:
            c = self.cfg
            z = np.zeros((c.jx, c.ix))
            self._rsolver = PIPSolver(c, np.zeros((5, c.jx, c.ix)), np.zeros((c.nvar_m, c.jx, c.ix)), z,
                                      prologue=False, tile=None)
            if self.replay_tile is not None:
                self._rsolver.set_tile(*self.replay_tile)
        return self._rsolver

    def run(self, U0: torch.Tensor, ac: torch.Tensor) -> torch.Tensor:
        """Differentiable run: U_T = run(U0, ac)."""
        return _Evolve.apply(U0, ac, self)

    def dt_history(self) -> np.ndarray:
        self._lib_done()
        return self.hist[: 2 * self.nsteps].view(-1, 2)[:, 0].cpu().numpy()

    def close(self):
        if self._rsolver is not None:
            self._rsolver.close()
            self._rsolver = None
        self.solver.close()


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, ac, runner: TwoFluidRun):
        U0c = U0.detach().contiguous()
        acc = ac.detach().contiguous()
        out, ckpts, hist = runner.forward_record(U0c, acc)
        ctx.runner = runner
        ctx.ckpts = ckpts
        ctx.run = runner.last_run
        # this call's own dt/cmax sequence: a later forward on the same
        # runner overwrites runner.hist, not this copy
        ctx.save_for_backward(acc, hist)
        return out

    @staticmethod
    def backward(ctx, gU):
        acc, hist = ctx.saved_tensors
        lam, lam_ac = ctx.runner.backward(gU, ctx.ckpts, acc, hist, run=ctx.run)
        return lam, lam_ac, None


def evolve(runner: TwoFluidRun, U0: torch.Tensor, ac: torch.Tensor) -> torch.Tensor:
    return runner.run(U0, ac)
