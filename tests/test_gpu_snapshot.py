"""GPU round trip of the reference-layout snapshots: a solver's output
(state downloaded from the device) restarts another solver bitwise, and the
restarted run continues the original one (ulp level: the restart runs the
prologue's rewrite round, as the reference's restart does).  With
ionisation/recombination the restart keeps the file's arb_heat."""
import os

import numpy as np
import pytest

from pip_amd import build, problems, snapshot
from pip_amd.config import PIPConfig
from pip_amd.solver import PIPSolver
from tests.util import h5_listing, snapshot_case

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(build.HDF5_PREFIX, "include", "hdf5.h")),
                                 reason="no HDF5 C library in this image")]


@pytest.mark.parametrize("case", ["orszag_tang", "ir_col3"])
def test_output_restart_continue(tmp_path, case):
    print(f"case snapshot restart {case}", flush=True)
    cfg, uh, um, ac = problems.orszag_tang(32)
    if case == "ir_col3":
        cfg = PIPConfig(**{**cfg.__dict__, "flag_ir": 3, "flag_ir_type": 0, "flag_col": 3})
        ac = None
    s = PIPSolver(cfg, uh, um, ac)
    s.advance(6)
    t6 = s.scalars()["time"]
    path = snapshot.write_solver(s, str(tmp_path), 1)
    assert os.path.exists(path)
    # the datasets the reference's output path creates for these flags
    ls = h5_listing(path)
    if ls is not None:
        want = snapshot_case("ir_3_type_0_col_3_later" if cfg.flag_ir else "orszag_tang_later_output")
        assert set(ls) == set(want["datasets"])
    snap = snapshot.read(str(tmp_path), 1, cfg, ac=True, rates=bool(cfg.flag_ir))
    h6, m6 = s.state()
    assert np.array_equal(snap["U_h"], h6) and np.array_equal(snap["U_m"], m6) and snap["time"] == t6
    assert np.array_equal(snap["ac"], s.ac_field())
    if cfg.flag_ir:
        gm_rec, gm_ion, arb = s.rates()
        assert np.array_equal(snap["gm_ion"], gm_ion) and np.array_equal(snap["arb_heat"], arb)
    z = np.zeros
    r = PIPSolver(cfg, z((5, cfg.jx, cfg.ix)), z((cfg.nvar_m, cfg.jx, cfg.ix)),
                  z((cfg.jx, cfg.ix)) if cfg.ac_field else None, prologue=False)
    snapshot.restart_solver(r, str(tmp_path), 1)
    assert r.scalars()["time"] == t6
    if cfg.flag_ir:
        # arb_heat is the file's, not fixed again from the restart state
        # (PIP_rot.f90:219); ion_pot starts at zero as in a restarted
        # reference run, so the first dt may differ from the continued run's
        assert np.array_equal(r.rates()[2], snap["arb_heat"])
        n, _ = r.advance(6)
        assert n == 6 and r.status() == 0
        assert np.array_equal(r.rates()[2], snap["arb_heat"])
    else:
        s.advance(6)
        r.advance(6)
        assert r.scalars()["time"] == pytest.approx(s.scalars()["time"], rel=1e-12)
        for a, b in zip(r.state(), s.state()):
            i = (slice(None), slice(2, -2), slice(2, -2))
            err = np.abs(a[i] - b[i]).sum() / np.abs(b[i]).sum()
            assert err < 1e-12, err
    s.close()
    r.close()
