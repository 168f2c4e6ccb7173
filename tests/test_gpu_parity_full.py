"""Parity of the HIP path with the CPU oracle at the configurations' real
sizes and step counts, bitwise:

  * 2-D two-fluid Orszag-Tang, 1024^2 periodic, 500 steps (the whole run);
  * 2-D two-fluid Kelvin-Helmholtz with set_IR case 3 (ion_pot / arb_heat),
    2048^2, x periodic, y mirror, the first 100 of its 1000 steps, and the
    same configuration at 384^2 for all 1000 steps;
  * 2-D Alfven-wave damping with a per-cell alpha_c field, 512^2, 200 steps
    (the adjoint configuration's forward run).

The oracle runs its elementwise loops on all host threads (its results do
not depend on the thread count).  Equality is asserted on every interior
cell of every variable, on the time-step sequence's end point (dt, time,
step count) and on the stop flags.

The same cases check the relaxed-arithmetic stage kernels (FMA contraction,
reciprocal-multiply quotients, hardware min/max) against the oracle after
100 steps: relative L1 <= 1e-10 for the variables of order one (mean
magnitude >= 5 % of the fluid's mean density) and <= 1e-5 for the others
(perturbation-born momenta and fields, B_z, psi, whose size is set by the
perturbation or by round-off).
"""
import sys
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from pip_amd import problems
from pip_amd.solver import PIPSolver
from tests.util import interior, relaxed_errors

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


RELAXED_STEPS = 100
RELAXED_TOL_BIG, RELAXED_TOL_SMALL = 1e-10, 1e-5


def _relaxed_check(cfg, uh, um, ac, o100):
    """The relaxed kernels after RELAXED_STEPS steps against the oracle's
    state at that step: relative L1 per variable."""
    s = PIPSolver(cfg, uh, um, ac, arith="relaxed")
    n, _ = s.advance(nsteps=RELAXED_STEPS)
    gh, gm = s.state()
    assert n == RELAXED_STEPS and s.status() == 0
    s.close()
    oh, om = o100
    big, small = relaxed_errors(gm, gh, om, oh, cfg)
    say(f"  relaxed after {RELAXED_STEPS} steps: rel L1 order-one variables {big:.2e}, others {small:.2e}")
    assert np.all(np.isfinite(gm)) and np.all(np.isfinite(gh))
    assert big <= RELAXED_TOL_BIG and small <= RELAXED_TOL_SMALL, (big, small)


def _bitwise_run(cfg, uh, um, ac, nsteps):
    t0 = time.perf_counter()
    s = PIPSolver(cfg, uh, um, ac)
    n, t = s.advance(nsteps=nsteps)
    gh, gm = s.state()
    gsc, gst = s.scalars(), s.status()
    s.close()
    say(f"  gpu: {n} steps in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    o = Oracle(cfg, uh, um, ac)
    st = o.run(nsteps=RELAXED_STEPS)
    _relaxed_check(cfg, uh, um, ac, o.state())
    if st == 0 and nsteps > RELAXED_STEPS:
        st = o.run(nsteps=nsteps - RELAXED_STEPS)
    oh, om = o.state()
    osc = o.scalars()
    say(f"  oracle: {osc['nt']} steps in {time.perf_counter() - t0:.1f} s")
    assert n == nsteps == osc["nt"]
    assert gst == st == 0
    assert gsc["dt"] == osc["dt"] and gsc["time"] == osc["time"] == t
    for name, g, r in (("U_m", gm, om), ("U_h", gh, oh)):
        gi, ri = interior(g, cfg), interior(r, cfg)
        diff = np.argwhere(gi != ri)
        assert diff.size == 0, f"{name}: {len(diff)} entries differ, first {diff[:4].tolist()}"
    say("  bitwise equal")


@pytest.mark.timeout(600)
def test_orszag_tang_1024_full_run():
    say("case orszag_tang n=1024 steps=500 vs oracle")
    cfg, uh, um, ac = problems.orszag_tang(1024)
    _bitwise_run(cfg, uh, um, ac, 500)


@pytest.mark.timeout(600)
def test_kelvin_helmholtz_ir_2048_100_steps():
    say("case kelvin_helmholtz n=2048 steps=100 vs oracle (IR case 3, mirror y)")
    cfg, uh, um, ac = problems.kelvin_helmholtz(2048)
    _bitwise_run(cfg, uh, um, ac, 100)


@pytest.mark.timeout(600)
def test_kelvin_helmholtz_ir_384_1000_steps():
    """The Kelvin-Helmholtz + IR configuration for its whole 1000 steps, at
    384^2 so that the oracle finishes in about a minute."""
    say("case kelvin_helmholtz n=384 steps=1000 vs oracle (IR case 3, mirror y)")
    cfg, uh, um, ac = problems.kelvin_helmholtz(384)
    _bitwise_run(cfg, uh, um, ac, 1000)


@pytest.mark.timeout(600)
def test_alfven_per_cell_alpha_512_full_run():
    say("case alfven n=512 steps=200 vs oracle (per-cell alpha_c)")
    cfg, uh, um, ac = problems.alfven_damping(512)
    _bitwise_run(cfg, uh, um, ac, 200)
