"""The Fortran driver (iso_c_binding module pipgpu) runs main.f90's loop
through the C-ABI; its result must equal the Python route bit for bit, and
the fused pip_advance path must agree with it."""
import os
import subprocess
import sys

import numpy as np
import pytest

from pip_amd import build, problems
from pip_amd.solver import PIPSolver

pytestmark = pytest.mark.gpu


def _run_driver(tmp_path, mode):
    if not os.path.exists(build.FBIN):
        pytest.skip("Fortran driver not built (flang missing)")
    out = tmp_path / f"out_{mode}.bin"
    args = [build.FBIN, "1024", str(out)] + (["fused"] if mode == "fused" else [])
    print("case fortran driver", mode, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    cfg, _, _, _ = problems.brio_wu(1024)
    raw = np.fromfile(out, dtype=np.float64)
    N = cfg.ix * cfg.jx
    uh = raw[: 5 * N].reshape(5, cfg.jx, cfg.ix)
    um = raw[5 * N:].reshape(8, cfg.jx, cfg.ix)
    return uh, um


def test_fortran_driver_matches_python_route(tmp_path):
    fh, fm = _run_driver(tmp_path, "loop")
    cfg, uh, um, ac = problems.brio_wu(1024)
    s = PIPSolver(cfg, uh, um, ac)
    while s.time_host < 0.1:
        s.step()
    ph, pm = s.state()
    assert np.array_equal(fh[:, :, 2:-2], ph[:, :, 2:-2])
    assert np.array_equal(fm[:, :, 2:-2], pm[:, :, 2:-2])
    gh, gm = _run_driver(tmp_path, "fused")
    assert np.array_equal(gm[:, :, 2:-2], fm[:, :, 2:-2])
    assert np.array_equal(gh[:, :, 2:-2], fh[:, :, 2:-2])
