"""The headline configuration's whole run against the oracle, bitwise:
Kelvin-Helmholtz with set_IR case 3, 2048^2, all 1000 steps.  The oracle
needs ~10 minutes for it on 16 host threads, so the test runs only when
PIP_LONG_PARITY=1 is set (it is skipped otherwise); its last result is
recorded with the measurements."""
import os
import sys
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from pip_amd import problems
from pip_amd.solver import PIPSolver
from tests.util import interior

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("PIP_LONG_PARITY") != "1",
                                 reason="long run: set PIP_LONG_PARITY=1")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(1100)
def test_kelvin_helmholtz_ir_2048_all_1000_steps():
    say("case kelvin_helmholtz n=2048 steps=1000 vs oracle (IR case 3, mirror y)")
    cfg, uh, um, ac = problems.kelvin_helmholtz(2048)
    s = PIPSolver(cfg, uh, um, ac)
    n, t = s.advance(nsteps=1000)
    gh, gm = s.state()
    gsc, gst = s.scalars(), s.status()
    s.close()
    o = Oracle(cfg, uh, um, ac)
    t0 = time.perf_counter()
    for k in range(10):
        st = o.run(nsteps=100)
        say(f"  oracle: {o.scalars()['nt']} steps, {time.perf_counter() - t0:.0f} s")
        assert st == 0
    oh, om = o.state()
    osc = o.scalars()
    assert n == 1000 == osc["nt"] and gst == 0
    assert gsc["dt"] == osc["dt"] and gsc["time"] == osc["time"] == t
    for name, g, r in (("U_m", gm, om), ("U_h", gh, oh)):
        diff = np.argwhere(interior(g, cfg) != interior(r, cfg))
        assert diff.size == 0, f"{name}: {len(diff)} entries differ, first {diff[:4].tolist()}"
    say("  bitwise equal after 1000 steps")
