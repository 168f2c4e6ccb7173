"""Snapshots in the reference's HDF5 layout (pip_amd.snapshot over
libpipio.so).

The dataset names, shapes and types follow the reference's output routines
(IO_rot.f90:140-374): coordinates xgrid/ygrid/zgrid/dx/dy/dz of length
ix/jx/kx, a scalar 'time', and every (ix,jx,kx) Fortran field as a (kx,jx,ix)
dataspace of IEEE little-endian doubles.  Which coefficient fields appear is
fixed by def_varfiles / save_varfiles (IO_rot.f90:231-330).  The listing is
taken with the HDF5 tools' h5dump -H; the round trip through the restart
reader is bitwise.
"""
import os

import numpy as np
import pytest

from pip_amd import build, problems, snapshot
from pip_amd.config import MARGIN, PIPConfig
from tests.util import SNAPSHOT_TABLE, h5_listing

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(build.HDF5_PREFIX, "include", "hdf5.h")),
                                reason="no HDF5 C library in this image")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not os.path.exists(build.IOLIB) or os.path.getmtime(build.IOLIB) < os.path.getmtime(
            os.path.join(build.CSRC, "pipio.cpp")):
        build.build_io()


def _listing(path):
    out = h5_listing(path)
    if out is None:
        pytest.skip("h5dump not installed")
    return out


PLASMA9 = ["ro_p", "mx_p", "my_p", "mz_p", "en_p", "bx", "by", "bz", "ps"]
NEUTRAL = ["ro_n", "mx_n", "my_n", "mz_n", "en_n"]
COORDS = ["xgrid", "ygrid", "zgrid", "dx", "dy", "dz"]


def test_orszag_tang_layout_and_round_trip(tmp_path):
    cfg, uh, um, _ = problems.orszag_tang(24)
    rng = np.random.default_rng(3)
    ac = np.full((cfg.jx, cfg.ix), cfg.col)
    xi = rng.uniform(0.1, 0.9, (cfg.jx, cfg.ix))
    path = snapshot.write(str(tmp_path), 7, cfg, uh, um, 0.125, ac=ac, xi=xi)
    assert os.path.basename(path) == "t0007.h5"
    ls = _listing(path)
    want = set(COORDS + ["time", "ac", "xi", "col"] + PLASMA9 + NEUTRAL)
    assert set(ls) == want
    assert ls["time"] == ("SCALAR", "H5T_IEEE_F64LE")
    for n in PLASMA9 + NEUTRAL + ["ac", "xi", "col"]:
        assert ls[n] == ((1, cfg.jx, cfg.ix), "H5T_IEEE_F64LE"), n
    assert ls["xgrid"][0] == (cfg.ix,) and ls["ygrid"][0] == (cfg.jx,) and ls["zgrid"][0] == (1,)
    snap = snapshot.read(str(tmp_path), 7, cfg, ac=True)
    assert snap["time"] == 0.125
    assert np.array_equal(snap["U_h"], uh) and np.array_equal(snap["U_m"], um)
    assert np.array_equal(snap["ac"], ac)


def test_one_dimensional_without_psi(tmp_path):
    cfg, uh, um, _ = problems.brio_wu(nx=64)
    assert cfg.nvar_m == 8
    path = snapshot.write(str(tmp_path), 0, cfg, uh, um, 0.0)
    ls = _listing(path)
    assert set(ls) == set(COORDS + ["time"] + PLASMA9[:8] + NEUTRAL)
    assert ls["ro_p"][0] == (1, 1, cfg.ix)
    snap = snapshot.read(str(tmp_path), 0, cfg)
    assert np.array_equal(snap["U_m"], um) and np.array_equal(snap["U_h"], uh)
    # unused directions: one point, d = 100 (Model_rot.f90:425-430)
    assert snap["dy"][0] == 100.0 and snap["dz"][0] == 100.0 and snap["y"][0] == 0.0


def test_save_keys_and_rate_fields(tmp_path):
    cfg, uh, um, _ = problems.orszag_tang(16)
    ir = PIPConfig(**{**cfg.__dict__, "flag_ir": 3, "flag_ir_type": 0})
    f = np.arange(ir.jx * ir.ix, dtype=np.float64).reshape(ir.jx, ir.ix)
    path = snapshot.write(str(tmp_path), 1, ir, uh, um, 1.0, ac=f, gm_ion=f + 1, gm_rec=f + 2, arb_heat=f + 3,
                          ps_sav=1, col_sav=1)
    ls = _listing(path)
    assert set(ls) == set(COORDS + ["time", "ac", "ion", "rec", "aheat"] + PLASMA9[:8] + NEUTRAL)
    # the reader needs 'ps' for nvar_m = 9, as the reference's restart does
    with pytest.raises(snapshot.PipioError, match="ps"):
        snapshot.read(str(tmp_path), 1, ir)
    path = snapshot.write(str(tmp_path), 2, ir, uh, um, 1.0, ac=f, gm_ion=f + 1, gm_rec=f + 2, arb_heat=f + 3)
    snap = snapshot.read(str(tmp_path), 2, ir, ac=True, rates=True)
    for k, v in (("ac", f), ("gm_ion", f + 1), ("gm_rec", f + 2), ("arb_heat", f + 3)):
        assert np.array_equal(snap[k], v), k
    # flag_IR = 1: 'ion' and 'rec' are created once (def_varfiles, then save_varfiles)
    ir1 = PIPConfig(**{**cfg.__dict__, "flag_ir": 1, "flag_ir_type": 1})
    ls = _listing(snapshot.write(str(tmp_path), 3, ir1, uh, um, 1.0, gm_ion=f, gm_rec=f))
    assert "ion" in ls and "rec" in ls and "aheat" not in ls


def test_grid_arrays_follow_set_coordinate():
    cfg = PIPConfig(ndim=2, nx=10, ny=6, xmin=-1.0, xmax=2.0, ymin=0.0, ymax=3.0)
    x, y, z, dx, dy, dz = snapshot.grid_arrays(cfg)
    assert dx[0] == 3.0 / 10 and dy[0] == 3.0 / 6 and dz[0] == 100.0 and z[0] == 0.0
    assert x[MARGIN] == -1.0 + 0.5 * dx[0] and y[MARGIN] == 0.0 + 0.5 * dy[0]
    # the recurrence: first differences are exactly dx from the origin cell up
    acc = x[MARGIN]
    for i in range(MARGIN + 1, cfg.ix):
        acc = acc + dx[0]
        assert x[i] == acc
    assert x[MARGIN - 1] == x[MARGIN] - dx[0] and x[0] == x[1] - dx[0]
    # and agree with the closed-form centres to rounding
    assert np.allclose(x, cfg.x_centers(), rtol=0, atol=1e-14)


def test_reader_errors(tmp_path):
    cfg, uh, um, _ = problems.orszag_tang(8)
    with pytest.raises(snapshot.PipioError, match="cannot open"):
        snapshot.read(str(tmp_path), 5, cfg)
    snapshot.write(str(tmp_path), 5, cfg, uh, um, 0.5)
    other = PIPConfig(**{**cfg.__dict__, "nx": 10})
    with pytest.raises(snapshot.PipioError, match="shape"):
        snapshot.read(str(tmp_path), 5, other)
    with pytest.raises(snapshot.PipioError, match="nout"):
        snapshot.write(str(tmp_path), 10000, cfg, uh, um, 0.5)


@pytest.mark.parametrize("case", SNAPSHOT_TABLE["cases"], ids=[c["name"] for c in SNAPSHOT_TABLE["cases"]])
def test_dataset_table_matches_the_reference_output_routine(tmp_path, case):
    """The datasets of one output, written through output_fields (the
    selection write_solver uses) with the case's flags and *_sav keys, are
    those the reference's output path creates (golden/snapshot_datasets.json,
    walked out of IO_rot.f90 by golden/make_snapshot_table.py); a name the
    reference creates twice is one dataset."""
    f = dict(case["flags"])
    base = problems.brio_wu(nx=16) if f["nvar_m"] == 8 else problems.orszag_tang(12)
    cfg, uh, um, _ = base
    cfg = PIPConfig(**{**cfg.__dict__, **{k: f[k] for k in ("flag_divb", "flag_col", "flag_ir", "flag_ir_type")}})
    assert cfg.nvar_m == f["nvar_m"]
    plane = np.ones((cfg.jx, cfg.ix))
    rates = (plane * 2, plane * 3, plane * 4) if cfg.flag_ir else None
    fields = snapshot.output_fields(cfg, plane, rates)
    sav = {k: f[k] for k in ("ac_sav", "xi_sav", "col_sav", "ion_sav", "rec_sav", "heat_sav", "ps_sav")}
    path = snapshot.write(str(tmp_path), case["nout"], cfg, uh, um, 0.5, **fields, **sav)
    ls = _listing(path)
    assert set(ls) == set(case["datasets"])
    for n in set(case["datasets"]) - set(COORDS + ["time"]):
        assert ls[n] == ((1, cfg.jx, cfg.ix), "H5T_IEEE_F64LE"), n
