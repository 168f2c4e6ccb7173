"""Batched ensemble (one persistent launch, optional hipGraph replay;
lockstep launches with every member spread over workgroups; or lockstep
chunks captured as a hipGraph) equals the plain per-member runs and the
oracle bit for bit."""
import numpy as np
import pytest

from oracle.oracle import Oracle

from pip_amd import problems
from pip_amd.ensemble import Ensemble
from pip_amd.solver import PIPSolver

pytestmark = pytest.mark.gpu


def _members(n, nx_cap=None):
    params = problems.ensemble_members(256, seed=7)[:n]
    out = []
    for p in params:
        if nx_cap:
            p = dict(p, nx=min(p["nx"], nx_cap))
        cfg, uh, um, _ = problems.ensemble_member(p)
        out.append((cfg, uh, um))
    return out


# every run is capped at MAXS steps (the members need about 3000 to reach
# t = 0.1), so a wrong dt cannot keep a launch running for long
MAXS = 20000


@pytest.mark.parametrize("mode,graph,tile", [("persistent", True, 256), ("persistent", False, 64),
                                             ("persistent", True, 512), ("lockstep", False, 256),
                                             ("lockstep", False, 64), ("lockstep", False, 512),
                                             ("graph", False, 256), ("graph", False, 64), ("graph", False, 512)])
def test_ensemble_matches_plain_runs(mode, graph, tile):
    print("case ensemble 10 members mode", mode, "graph", graph, "tile", tile, flush=True)
    mem = _members(10)
    ens = Ensemble(mem, tend=0.1, max_steps=MAXS, tile=tile)
    ens.set_mode(mode)
    ms = ens.run(graph=graph)
    print("ensemble ms", ms, flush=True)
    for k, (cfg, uh, um) in enumerate(mem):
        s = PIPSolver(cfg, uh, um)
        n, t = s.advance(nsteps=MAXS, tend=0.1)
        r = ens.result(k)
        assert r["status"] == 0
        assert r["nt"] == n and r["time"] == t, (k, r, n, t)
        eh, em = ens.state(k)
        ph, pm = s.state()
        assert np.array_equal(em[:, :, 2:-2], pm[:, :, 2:-2]), k
        assert np.array_equal(eh[:, :, 2:-2], ph[:, :, 2:-2]), k
        s.close()


@pytest.mark.parametrize("mode", ["persistent", "lockstep", "graph"])
def test_ensemble_rerun_after_reset_is_identical(mode):
    """A rerun from the initial states repeats the first run bit for bit; in
    graph mode the rerun replays the cached graphs of the shrinking member
    sets the first run captured."""
    print("case ensemble reset/rerun", mode, flush=True)
    mem = _members(6, nx_cap=600)
    ens = Ensemble(mem, tend=0.05, max_steps=MAXS)
    ens.set_mode(mode)
    ens.run(graph=True)
    a = [(ens.result(k), ens.state(k)) for k in range(6)]
    ens.reset()
    ens.run(graph=True)
    for k in range(6):
        r, b = ens.result(k), ens.state(k)
        assert r == a[k][0]
        assert np.array_equal(a[k][1][0], b[0]) and np.array_equal(a[k][1][1], b[1])


@pytest.mark.parametrize("cap", [150, 128, 64])
def test_lockstep_equals_persistent_at_the_step_cap(cap):
    """Lockstep and persistent launches give the same states (ghosts
    included), times, step counts and statuses when the members stop at the
    step cap inside a 64-step chunk (150: the next step's stage 1 applies the
    step end) or at a chunk's last step (128, 64: the chunk's closing step
    end), and again after a reset."""
    print("case ensemble lockstep vs persistent, 12 members, cap", cap, flush=True)
    mem = _members(12, nx_cap=1500)
    ens = Ensemble(mem, tend=0.1, max_steps=cap, tile=128)
    ens.run(graph=False)
    ref = [(ens.result(k), ens.state(k)) for k in range(12)]
    assert all(r["nt"] == cap for r, _ in ref)
    for mode in ("lockstep", "graph"):
        ens.set_mode(mode)
        for _ in range(2):
            ens.reset()
            ens.run()
            for k in range(12):
                r, (uh, um) = ref[k]
                assert ens.result(k) == r, (mode, k)
                eh, em = ens.state(k)
                assert np.array_equal(eh, uh) and np.array_equal(em, um), (mode, k)


def test_lockstep_and_persistent_agree_on_48_members():
    """48 members of uneven size (chunks drop members at different steps):
    lockstep at width 128 against the persistent launch at width 256, every
    state (ghosts included), time, step count and status equal."""
    print("case ensemble 48 members lockstep 128 vs persistent 256, t_end 0.02", flush=True)
    mem = _members(48, nx_cap=800)
    ens = Ensemble(mem, tend=0.02, max_steps=MAXS, tile=256)
    ens.run(graph=True)
    ref = [(ens.result(k), ens.state(k)) for k in range(48)]
    assert len({r["nt"] for r, _ in ref}) > 10  # members stop at many different steps
    for mode in ("lockstep", "graph"):
        ens.set_mode(mode)
        ens.set_tile(128)
        ens.reset()
        ens.run()
        for k in range(48):
            r, (uh, um) = ref[k]
            assert ens.result(k) == r, (mode, k)
            eh, em = ens.state(k)
            assert np.array_equal(eh, uh) and np.array_equal(em, um), (mode, k)


@pytest.mark.timeout(600)
def test_ensemble_members_match_the_oracle():
    """Members of the 256-member ensemble (seed 7) run to t = 0.1 inside a
    24-member ensemble, in all three launch modes, against the oracle's runs of the
    same members: the stiffest (alpha_c = 9.8e3), the largest (nx = 4089),
    the smallest (nx = 260) and two others -- every interior cell, the step
    count and the end time bitwise."""
    print("case ensemble vs oracle: members 175, 98, 212, 3, 100 in a 24-member ensemble", flush=True)
    params = problems.ensemble_members(256, seed=7)
    check = [175, 98, 212, 3, 100]
    assert params[175]["col"] > 9.7e3 and params[98]["nx"] == max(p["nx"] for p in params)
    pick = check + [k for k in range(19)]
    mem = [problems.ensemble_member(params[k])[:3] for k in pick]
    ref = []
    for k in range(len(check)):
        cfg, uh, um = mem[k]
        o = Oracle(cfg, uh, um)
        assert o.run(nsteps=MAXS, tend=0.1) == 0
        ref.append((o.scalars(), o.state()))
        print(f"  oracle member {check[k]}: nx {cfg.nx}, {ref[-1][0]['nt']} steps", flush=True)
    ens = Ensemble(mem, tend=0.1, max_steps=MAXS, tile=256)
    for mode in ("persistent", "lockstep", "graph"):
        ens.set_mode(mode)
        ens.reset()
        ens.run(graph=True)
        for k in range(len(check)):
            sc, (oh, om) = ref[k]
            r = ens.result(k)
            assert r["status"] == 0 and r["nt"] == sc["nt"] and r["time"] == sc["time"], (mode, check[k], r, sc)
            eh, em = ens.state(k)
            assert np.array_equal(em[:, :, 2:-2], om[:, :, 2:-2]), (mode, check[k])
            assert np.array_equal(eh[:, :, 2:-2], oh[:, :, 2:-2]), (mode, check[k])
        print(f"  {mode}: bitwise equal", flush=True)
    ens.close()


def _members_2d():
    """Uneven 2-D two-fluid members (Orszag-Tang initial states, psi) of
    sizes that are not multiples of the 16x15 tile, sweeping the neutral
    fraction and the collision coefficient."""
    out = []
    for n, xi, col in ((24, 0.9, 1.0e3), (40, 0.5, 30.0), (33, 0.2, 3.0e3), (64, 0.7, 100.0),
                       (17, 0.95, 10.0), (50, 0.3, 1.0e4)):
        cfg, uh, um, _ = problems.orszag_tang(n, xi_n=xi, col=col)
        out.append((cfg, uh, um))
    return out


@pytest.mark.parametrize("mode", ["lockstep", "graph"])
def test_2d_ensemble_matches_plain_and_oracle(mode):
    """2-D members (psi, constant alpha_c) of six sizes run 90 steps in one
    ensemble: every member equals its plain pip_advance run bit for bit
    (ghosts excluded), time and step count included; two members also
    equal the oracle's run."""
    print("case 2-D ensemble, 6 members, 90 steps, mode", mode, flush=True)
    mem = _members_2d()
    ens = Ensemble(mem, tend=1.0e30, max_steps=90)
    ens.set_mode(mode)
    ens.run()
    for k, (cfg, uh, um) in enumerate(mem):
        s = PIPSolver(cfg, uh, um)
        n, t = s.advance(nsteps=90)
        r = ens.result(k)
        assert r["status"] == 0 and r["nt"] == n == 90 and r["time"] == t, (k, r, n, t)
        eh, em = ens.state(k)
        ph, pm = s.state()
        assert np.array_equal(em[:, 2:-2, 2:-2], pm[:, 2:-2, 2:-2]), k
        assert np.array_equal(eh[:, 2:-2, 2:-2], ph[:, 2:-2, 2:-2]), k
        s.close()
        if k in (1, 2):
            o = Oracle(cfg, uh, um)
            o.run(nsteps=90)
            oh, om = o.state()
            assert np.array_equal(em[:, 2:-2, 2:-2], om[:, 2:-2, 2:-2]), k
            assert np.array_equal(eh[:, 2:-2, 2:-2], oh[:, 2:-2, 2:-2]), k
    ens.close()


def test_2d_ensemble_rejects_the_persistent_mode():
    print("case 2-D ensemble persistent mode rejected", flush=True)
    from pip_amd._lib import PipError
    ens = Ensemble(_members_2d()[:2], tend=1.0, max_steps=5)
    ens.set_mode("persistent")
    with pytest.raises(PipError, match="2-D ensembles run in lockstep"):
        ens.run()
    ens.close()


def _stop_members(twod):
    """Members of which one plants a NaN in the magnetic field and one takes
    a collision coefficient so large that dt falls below its floor
    (IO_rot.f90:625-674); the others run normally."""
    if twod:
        mem = _members_2d()[:4]
        cfg, uh, um = mem[1]
        um = um.copy()
        um[7, 9, 14] = np.nan
        mem[1] = (cfg, uh, um)
        cfg, uh, um, _ = problems.orszag_tang(24, col=1.0e14)
        mem[3] = (cfg, uh, um)
    else:
        mem = _members(6, nx_cap=700)
        cfg, uh, um = mem[1]
        um = um.copy()
        um[6, 0, 30] = np.nan
        mem[1] = (cfg, uh, um)
        p = dict(problems.ensemble_members(256, seed=7)[4], nx=300, col=1.0e13)
        mem[4] = problems.ensemble_member(p)[:3]
    return mem


@pytest.mark.parametrize("mode,twod", [("persistent", False), ("lockstep", False), ("graph", False),
                                       ("lockstep", True), ("graph", True)])
def test_members_stop_like_their_plain_runs(mode, twod):
    """A member whose state holds a NaN stops after its first step with the
    NaN flag, one whose dt falls below the floor stops with the dt flag, as
    their plain pip_advance runs do; the other members run on, bitwise equal
    to their plain runs."""
    print("case ensemble stop_sim members, mode", mode, "2-D" if twod else "1-D", flush=True)
    mem = _stop_members(twod)
    steps = 60 if twod else MAXS
    ens = Ensemble(mem, tend=1.0e30 if twod else 0.02, max_steps=steps)
    ens.set_mode(mode)
    ens.run()
    stops = set()
    for k, (cfg, uh, um) in enumerate(mem):
        s = PIPSolver(cfg, uh, um)
        n, t = s.advance(nsteps=steps, tend=1.0e30 if twod else 0.02)
        r = ens.result(k)
        assert r["status"] == s.status(), (k, r, s.status())
        assert r["nt"] == n, (k, r, n)
        if not r["status"] & 2:
            # (the dt of a state holding a NaN is not defined by the
            # reference: only the stop and its step are compared)
            assert r["time"] == t, (k, r, t)
        if r["status"]:
            stops.add(r["status"] & 3)
        else:
            eh, em = ens.state(k)
            ph, pm = s.state()
            i = (slice(None), slice(2, -2), slice(2, -2)) if twod else (slice(None), slice(None), slice(2, -2))
            assert np.array_equal(em[i], pm[i]) and np.array_equal(eh[i], ph[i]), k
        s.close()
    ens.close()
    assert stops == {1, 2}
